// Tick engine: the batched driver of every PSPG and weakly compressible request.
//
// The PSPG step restates PicardAlgo::solve (nonLinearAlgo/PicardAlgo.cpp:31-95) with the
// PSPG callbacks of MomContEquationPSPG.inl:262-372 as a per-request state machine:
//   prepare (save nodes, q = qPrev) -> assemble -> [ solve -> update positions ->
//   re-assemble -> residual -> test ]*  -> commit, or restore on failure.
// The linear solve replaces SparseLU (MomContEquationPSPG.inl:281-290) with BiCGSTAB +
// block Jacobi in FP64, or FP64 iterative refinement around FP32 BiCGSTAB.  Every
// decision is taken on the host from the slot scalars read back at the end of a tick;
// every operation is one launch over all requests that need it in that tick.
#include "runtime.hpp"

#include <algorithm>
#include <deque>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <type_traits>

namespace pfem {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

namespace {

constexpr int kElemCap = 8192;  // workgroups per case for element / node passes

// workgroups per case in row launches (<= kMaxCaseBlocks, the partials capacity);
// PFEM_ROW_BLOCKS overrides it (tuning)
const int kRowCap = [] {
    const char* e = std::getenv("PFEM_ROW_BLOCKS");
    const int v = e ? std::atoi(e) : 0;
    return (v >= 64 && v <= kMaxCaseBlocks) ? v : 1024;
}();

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int lane_index(int lanes) { return lanes == 4 ? 0 : (lanes == 8 ? 1 : 2); }
constexpr int kLanes[3] = {4, 8, 16};


enum : int { W_NONE = 0, W_K64 = 1, W_K32 = 2, W_IR = 3, W_POST = 4, W_RESID = 5 };

// residual kind of a request: its own (pfem_residual) or the case's
int res_kind_of(const Req* r) { return r->res_kind >= 0 ? r->res_kind : r->c->d.residual; }

enum Tab : int {
    T_COPY, T_UPD, T_ASM2, T_ASM3, T_BC, T_RESU, T_IRU,
    T_RES, T_IRR = T_RES + 3, T_K64 = T_IRR + 3, T_K32 = T_K64 + 3, T_F64 = T_K32 + 3, T_F32 = T_F64 + 3,
    T_WCN = T_F32 + 3, T_WCE2 = T_WCN + 2, T_WCE3 = T_WCE2 + 2, T_CFE2 = T_WCE3 + 2, T_CFE3 = T_CFE2 + 2,
    T_WEL2 = T_CFE3 + 2, T_WEL3 = T_WEL2 + 2,  // WC continuity element pass (one lane per element)
    T_AMG = T_WEL3 + 2,  // multigrid tables: FP64 hierarchies, then FP32 ones (A_N each)
};

// Multigrid tables of one hierarchy type, relative to T_AMG + type * A_N: setup per level
// (Galerkin, block Jacobi), coarsest dense inverse; V-cycle per level (node ops, SpMV ops,
// restriction), coarsest solve.
enum AmgTab : int {
    A_G = 0, A_B = A_G + kAmgMaxLevels, A_D = A_B + kAmgMaxLevels, A_AP = A_D + 1,
    A_VN = A_AP + kAmgMaxLevels, A_VS = A_VN + kAmgMaxLevels, A_VR = A_VS + kAmgMaxLevels,
    A_VC = A_VR + kAmgMaxLevels, A_N = A_VC + 1
};
constexpr int T_NTAB = T_AMG + 2 * A_N;

void append(std::vector<BlockTask>& t, int slot, long long work, int cap, int pad = 0) {
    long long nb = (work + kBlock - 1) / kBlock;
    nb = std::max<long long>(1, std::min<long long>(nb, cap));
    for (int b = 0; b < nb; ++b) t.push_back(BlockTask{slot, b, (int)nb, pad});
}

void append_rows(std::vector<BlockTask>& t, int slot, long long n, int nsl) {
    const int nb = row_blocks(n, nsl);
    for (int b = 0; b < nb; ++b) t.push_back(BlockTask{slot, b, nb, 0});
}

template <class T>
NodeMat<T> node_mat(const Pattern& P, int dim, const T* vals) {
    NodeMat<T> A{};
    A.N = P.N;
    A.nf = dim + 1;
    A.mom_nnz = P.mom_nnz;
    A.info = P.node_info.as<int4>();
    A.nbr_ptr = P.nbr_ptr.as<int>();
    A.nbr_idx = P.nbr_idx.as<int>();
    A.vals = vals;
    return A;
}

template <class T>
EllMatT<T> ell_mat(const Pattern& P, int dim, const T* vals) {
    EllMatT<T> E{};
    E.N = P.N;
    E.nf = dim + 1;
    E.nsl = P.ell_nsl;
    E.soff = P.ell_soff.as<int>();
    E.idx = P.ell_idx.as<int>();
    E.vals = vals;
    return E;
}

template <class T>
void pinned_alloc(T*& p, size_t count) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    PFEM_HIP(hipHostMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1), hipHostMallocDefault));
    std::memset((void*)p, 0, sizeof(T) * std::max<size_t>(count, 1));
}

int pow2ceil(int v) {
    int p = 1;
    while (p < v && p < (1 << 20)) p <<= 1;
    return p;
}

}  // namespace

int row_blocks(long long n, int nsl) {
    long long nb = std::max<long long>(1, std::min<long long>((n + kBlock - 1) / kBlock, kRowCap));
    if (nsl > 0) {
        // The ELL passes give each wavefront one 64-node slice at a time, over kXcd segments
        // of up to ceil(nsl / kXcd) slices: enough wavefronts per segment that every one takes
        // the same number of slices (one wavefront left with an extra slice runs it alone
        // after the others have finished).
        const long long seg = (nsl + kXcd - 1) / kXcd;
        const long long most = (long long)kMaxCaseBlocks / kXcd * kWaves;  // wavefronts per segment
        const long long rounds = (seg + most - 1) / most;
        const long long waves = (seg + rounds - 1) / rounds;
        nb = std::max<long long>(nb, kXcd * ((waves + kWaves - 1) / kWaves));
    }
    return (int)nb;
}

// lanes per field of the node-block SpMV: the node-graph degree (incl. the node itself)
// is ~7 for 2D Kuhn triangulations and ~15 for 3D Kuhn tetrahedralisations
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

// ---------------------------------------------------------------------------------------
// Workspace and arena
// ---------------------------------------------------------------------------------------
namespace {
struct WsLayout {
    size_t vals, b, binv, x, x_prev, pos_save, kv64, vals_ell, vals32_ell, vals32, binv32, kv32, erec, esv, total;
    size_t amg_w0, amg_t0, amg_dense;
    size_t amg_lv[kAmgMaxLevels][6];
    size_t amg_ap[kAmgMaxLevels];  // A P values of level l (rows) against level l+1
};
// The multigrid hierarchy runs in FP32 for PFEM_PREC_MIXED 3D cases (under FP64 BiCGSTAB):
// the 3D iteration counts did not change, and the all-FP32 3D hierarchy served the stream
// 8 % faster than FP32 operators with FP64 vectors.  2D hierarchies keep FP64 vectors: an
// FP32 V-cycle raised the 2D BiCGSTAB iteration count by 20-40 % (their level 0 reads FP32
// values instead, amg_l0_fp32).
// A hierarchy wholly in FP32 (values, inverses and vectors of every level): only for 3D
// PFEM_PREC_MIXED cases set up with PFEM_AMG_FP32=1.  On the 512-request serving stream it
// takes more BiCGSTAB iterations than an FP64 hierarchy with FP32 level-0 values (below):
// 24,313 against 23,519, and 210.6 against 220.4 steps/s; in 2D 31,053 iterations.
bool amg_fp32(const pfem_case* c) {
    return c->amg_all_fp32 && c->d.precision == PFEM_PREC_MIXED && c->amg.levels > 1 && c->dim == 3;
}

// PFEM_PREC_MIXED cases otherwise keep the FP64 hierarchy but read FP32 values at level 0:
// the residual t = r - A w (an FP32 ELL copy of the system) and the post-smoothing product
// (A P) z (A P held in FP32), with FP64 vectors and sums.  The iteration counts do not
// rise, and with non-temporal loads FP32 values nearly halve those kernels' time
// (measured on gfx950, ELL SpMV of a 900k-node 2D system: 91 -> 52 us).
bool amg_l0_fp32(const pfem_case* c) {
    return c->d.precision == PFEM_PREC_MIXED && c->amg.levels > 1 && !amg_fp32(c);
}

WsLayout ws_layout(const pfem_case* c) {
    WsLayout L{};
    const size_t n = (size_t)c->n, nnz = (size_t)c->pat.nnz, nf = (size_t)c->dim + 1, N = (size_t)c->N;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += align256(bytes);
        return at;
    };
    L.vals = take(8 * nnz);
    L.b = take(8 * n);
    L.binv = take(8 * nf * nf * N);
    L.x = take(8 * n);
    L.x_prev = take(8 * n);
    L.pos_save = take(8 * (size_t)c->dim * N);
    L.kv64 = take(8 * 8 * n);
    L.vals_ell = take(8 * nf * nf * kEllSlice * (size_t)c->pat.ell_ranks);
    L.erec = take(8 * 16 * (size_t)c->E);  // assembly element records (pspg.hip)
    L.esv = take(8 * 4 * (size_t)c->E);
    const bool ir = c->d.precision == PFEM_PREC_MIXED_IR;
    if (ir || amg_fp32(c)) {  // FP32 copies of the values and of the block-Jacobi inverse
        L.vals32 = take(4 * nnz);
        L.binv32 = take(4 * nf * nf * N);
    }
    if (ir) L.kv32 = take(4 * 10 * n);
    if (ir || amg_fp32(c) || amg_l0_fp32(c)) L.vals32_ell = take(4 * nf * nf * kEllSlice * (size_t)c->pat.ell_ranks);
    const AmgHierarchy& H = c->amg;
    if (H.levels > 1) {
        const size_t es = amg_fp32(c) ? 4 : 8;
        L.amg_w0 = take(es * n);
        L.amg_t0 = take(es * n);
        for (int l = 1; l < H.levels; ++l) {
            const AmgLevel& lv = H.lev[l];
            const size_t nl = (size_t)nf * lv.N;
            L.amg_lv[l][0] = take(es * (size_t)lv.nnz);
            L.amg_lv[l][1] = take(es * nf * nf * (size_t)lv.N);
            for (int k = 2; k < 6; ++k) L.amg_lv[l][k] = take(es * nl);
        }
        for (int l = 0; l + 1 < H.levels; ++l) {
            const size_t ea = (l == 0 && amg_l0_fp32(c)) ? 4 : es;
            L.amg_ap[l] = take(ea * nf * nf * kEllSlice * (size_t)H.lev[l + 1].ap_ranks);
        }
        L.amg_dense = take(8 * (size_t)H.coarse_n * H.coarse_n);  // FP64 in both types
    }
    L.total = o;
    return L;
}
}  // namespace

// The case's surface-tension facets as the kernels see them (fst.F == 0: no term).  all_nodes:
// the WC rule (a facet acts when every node is on the free surface), else the PSPG one (any node).
FstView fst_view(const pfem_case* c, bool all_nodes) {
    FstView v;
    const bool bq = c->d.kind == PFEM_KIND_BOUSSINESQ;  // other problems read gamma alone
    const bool on = !(c->d.gamma < 1e-15 && (!bq || c->d.dgamma_dt < 1e-15));
    if (!on || c->fst.F == 0) return v;
    v.F = c->fst.F;
    v.all_nodes = all_nodes ? 1 : 0;
    v.nodes = c->fst.nodes.as<int>();
    v.elem = c->fst.elem.as<int>();
    v.ptr = c->fst.ptr.as<int>();
    v.ent = c->fst.ent.as<unsigned>();
    v.gamma = c->d.gamma;
    if (c->d.kind == PFEM_KIND_BOUSSINESQ) {
        v.T = c->heat.T.as<double>();
        v.dgdT = c->d.dgamma_dt;
        v.t_ref = c->d.t_ref;
    }
    return v;
}

size_t workspace_bytes(const pfem_case* c) {
    if (!pspg_like(c) || !c->has_mesh) return 0;
    return ws_layout(c).total;
}

void carve_workspace(pfem_case* c, char* base) {
    const WsLayout L = ws_layout(c);
    Workspace& w = c->ws;
    w.base = base;
    w.bytes = L.total;
    w.vals = reinterpret_cast<double*>(base + L.vals);
    w.b = reinterpret_cast<double*>(base + L.b);
    w.binv = reinterpret_cast<double*>(base + L.binv);
    w.x = reinterpret_cast<double*>(base + L.x);
    w.x_prev = reinterpret_cast<double*>(base + L.x_prev);
    w.pos_save = reinterpret_cast<double*>(base + L.pos_save);
    w.kv64 = reinterpret_cast<double*>(base + L.kv64);
    w.vals_ell = reinterpret_cast<double*>(base + L.vals_ell);
    w.erec = reinterpret_cast<double*>(base + L.erec);
    w.esv = reinterpret_cast<double*>(base + L.esv);
    const bool ir = c->d.precision == PFEM_PREC_MIXED_IR;
    const bool v32 = ir || amg_fp32(c);
    w.vals32 = v32 ? reinterpret_cast<float*>(base + L.vals32) : nullptr;
    w.binv32 = v32 ? reinterpret_cast<float*>(base + L.binv32) : nullptr;
    w.kv32 = ir ? reinterpret_cast<float*>(base + L.kv32) : nullptr;
    w.vals32_ell = (ir || amg_fp32(c) || amg_l0_fp32(c)) ? reinterpret_cast<float*>(base + L.vals32_ell) : nullptr;
    const AmgHierarchy& H = c->amg;
    w.amg_w0 = w.amg_t0 = w.amg_dense = nullptr;
    for (auto& lv : w.amg_lv)
        for (auto& p : lv) p = nullptr;
    for (auto& p : w.amg_ap) p = nullptr;
    if (H.levels > 1) {
        w.amg_w0 = base + L.amg_w0;
        w.amg_t0 = base + L.amg_t0;
        w.amg_dense = base + L.amg_dense;
        for (int l = 1; l < H.levels; ++l)
            for (int k = 0; k < 6; ++k) w.amg_lv[l][k] = base + L.amg_lv[l][k];
        for (int l = 0; l + 1 < H.levels; ++l) w.amg_ap[l] = base + L.amg_ap[l];
    }
}

void own_workspace(pfem_case* c, hipStream_t s) {
    if (c->ws.base || !pspg_like(c) || !c->has_mesh) return;
    const size_t wb = workspace_bytes(c);
    c->ws.own.reserve(wb);
    carve_workspace(c, c->ws.own.as<char>());
    PFEM_HIP(hipMemcpyAsync(c->ws.x, c->q_prev.p, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToDevice, s));
}

void Arena::init(size_t bytes) {
    bytes = align256(std::max<size_t>(bytes, 256));
    mem_.reserve(bytes);
    cap_ = bytes;
    used_ = peak_ = 0;
    free_.clear();
    free_[0] = bytes;
}

bool Arena::alloc(size_t bytes, size_t& off) {
    bytes = align256(std::max<size_t>(bytes, 256));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        off = it->first;
        const size_t rem = it->second - bytes;
        free_.erase(it);
        if (rem) free_[off + bytes] = rem;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
    }
    return false;
}

void Arena::release(size_t off, size_t bytes) {
    bytes = align256(std::max<size_t>(bytes, 256));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// ---------------------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------------------
struct Engine::Impl {
    int device = 0;
    hipStream_t s = nullptr;
    int S = 0;
    Arena* arena = nullptr;
    DevBuf d_pv, d_pic, d_k64, d_k32, d_ref, d_wc64, d_wc32, d_scal, d_tab, d_amg64, d_amg32;
    AmgView<double>* h_amg64 = nullptr;  // [S * kAmgMaxLevels]
    AmgView<float>* h_amg32 = nullptr;
    double amg_omega = 0.9;  // fewest BiCGSTAB iterations among 0.5 / 0.7 / 0.9 on 2D and 3D PSPG systems
    double amg_alpha = 1.6;  // fewest iterations among 1.0 ... 2.2 on 2D (20k, 100k nodes) and 3D systems
    // keep the coarse operators over a step's Picard re-assemblies: the same BiCGSTAB
    // iteration counts on the serving stream, 8 % more throughput
    bool amg_reuse = true;
    PspgView* h_pv = nullptr;
    PicardView* h_pic = nullptr;
    KrylovView<double>* h_k64 = nullptr;
    KrylovView<float>* h_k32 = nullptr;
    RefineView* h_ref = nullptr;
    WcView<double>* h_wc64 = nullptr;
    WcView<float>* h_wc32 = nullptr;
    SlotScal* h_scal = nullptr;   // readback
    SlotScal* h_reset = nullptr;  // staging of host-written slot scalars
    BlockTask* h_tab = nullptr;
    size_t tab_cap = 0;
    std::vector<int> free_slots;
    int par = 0;
    bool prof = false;
    hipEvent_t ev[8] = {};
    hipEvent_t pev[128] = {};
    // the chunk of BiCGSTAB iterations (with the V-cycles between their phases) captured in a
    // HIP graph and replayed whenever a tick repeats it exactly: the same work tables at the
    // same device addresses, the same views, chunk length and iteration parity.  A chunk is
    // captured the second time its key comes up among the last kGraphSeen eager chunks (a table
    // set that changes every tick, as a busy server's does, is launched kernel by kernel and
    // never pays for a capture).
    bool graphs = false;
    struct Captured {
        std::vector<long long> key;
        hipGraphExec_t exec = nullptr;
    };
    std::deque<Captured> gcache;  // the most recent captures, at most kGraphCache
    static constexpr size_t kGraphCache = 8;
    std::deque<std::vector<long long>> seen;  // keys of the recent eager chunks, at most kGraphSeen
    static constexpr size_t kGraphSeen = 16;
    long long graph_launches = 0, graph_captures = 0;

    ~Impl() {
        for (auto& g : gcache)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : pev)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)h_pv, (void*)h_pic, (void*)h_k64, (void*)h_k32, (void*)h_ref, (void*)h_wc64,
                        (void*)h_wc32, (void*)h_scal, (void*)h_reset, (void*)h_tab, (void*)h_amg64,
                        (void*)h_amg32})
            if (p) (void)hipHostFree(p);
    }

    void alloc_slots(int n) {
        S = n;
        d_pv.reserve(sizeof(PspgView) * n);
        d_pic.reserve(sizeof(PicardView) * n);
        d_k64.reserve(sizeof(KrylovView<double>) * n);
        d_k32.reserve(sizeof(KrylovView<float>) * n);
        d_ref.reserve(sizeof(RefineView) * n);
        d_wc64.reserve(sizeof(WcView<double>) * n);
        d_wc32.reserve(sizeof(WcView<float>) * n);
        d_scal.reserve(sizeof(SlotScal) * n);
        d_amg64.reserve(sizeof(AmgView<double>) * n * kAmgMaxLevels);
        d_amg32.reserve(sizeof(AmgView<float>) * n * kAmgMaxLevels);
        PFEM_HIP(hipMemsetAsync(d_scal.p, 0, sizeof(SlotScal) * n, s));
        pinned_alloc(h_amg64, (size_t)n * kAmgMaxLevels);
        pinned_alloc(h_amg32, (size_t)n * kAmgMaxLevels);
        pinned_alloc(h_pv, n);
        pinned_alloc(h_pic, n);
        pinned_alloc(h_k64, n);
        pinned_alloc(h_k32, n);
        pinned_alloc(h_ref, n);
        pinned_alloc(h_wc64, n);
        pinned_alloc(h_wc32, n);
        pinned_alloc(h_scal, n);
        pinned_alloc(h_reset, n);
        PFEM_HIP(hipStreamSynchronize(s));
        free_slots.clear();
        for (int i = n - 1; i >= 0; --i) free_slots.push_back(i);
    }

    SlotScal* dslot(int slot) const { return d_scal.as<SlotScal>() + slot; }

    template <class V>
    void upload(DevBuf& d, V* h, int slot) {
        PFEM_HIP(hipMemcpyAsync(d.as<V>() + slot, h + slot, sizeof(V), hipMemcpyHostToDevice, s));
    }

    void load_views(Req* r);
    template <class T>
    void load_amg(pfem_case* c, int slot, const KrylovView<double>& k, AmgView<T>* h, DevBuf& d,
                  const NodeMat<T>& A0, const T* binv0);
    void reset_kscal(int slot, int which, double rtol, int maxit, double atol = 0.0);
    void restart_kscal(int slot, int which);
};

Engine::Engine(int device, hipStream_t s, int slots, Arena* arena) : p_(new Impl) {
    p_->device = device;
    p_->s = s;
    p_->arena = arena;
    for (auto& e : p_->ev) PFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    // PFEM_AMG_OMEGA overrides the multigrid smoother damping (tuning)
    if (const char* e = std::getenv("PFEM_AMG_OMEGA")) {
        const double w = std::atof(e);
        if (w > 0.0 && w <= 1.5) p_->amg_omega = w;
    }
    if (const char* e = std::getenv("PFEM_AMG_REUSE")) p_->amg_reuse = std::atoi(e) != 0;
    // PFEM_GRAPH=1 replays repeated BiCGSTAB iteration chunks from captured HIP graphs
    if (const char* e = std::getenv("PFEM_GRAPH")) p_->graphs = std::atoi(e) != 0;
    // PFEM_AMG_ALPHA overrides the coarse-correction scaling (tuning)
    if (const char* e = std::getenv("PFEM_AMG_ALPHA")) {
        const double a = std::atof(e);
        if (a > 0.0 && a <= 3.0) p_->amg_alpha = a;
    }
    p_->alloc_slots(std::max(1, slots));
}

Engine::~Engine() = default;

void Engine::ensure_slots(int s) {
    PFEM_REQUIRE(idle(), E_STATE, "engine resized while busy");
    if (s > p_->S) p_->alloc_slots(s);
}

void Engine::enable_profiling(bool on) {
    if (on && !p_->pev[0])  // timing-only events: no system-scope fence on record
        for (auto& e : p_->pev) PFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    p_->prof = on;
    if (on) prof_ms[0] = prof_ms[1] = 0, prof_count[0] = prof_count[1] = 0;
}

void Engine::enqueue(Req* r) { waitq_.push_back(r); }

// Views of the request's case, pointing the scalars at the slot, uploaded to the slot.
void Engine::Impl::load_views(Req* r) {
    pfem_case* c = r->c;
    const int slot = r->slot;
    const int N = c->N, dim = c->dim;
    const long long n = c->n;
    SlotScal* ds = dslot(slot);
    if (pspg_like(c)) {
        const Pattern& P = c->pat;
        const Workspace& w = c->ws;
        const bool mixed = c->d.precision == PFEM_PREC_MIXED_IR;  // FP32 inner solves
        PspgView& pv = h_pv[slot];
        pv = PspgView{};
        pv.dim = dim;
        pv.N = N;
        pv.E = c->E;
        pv.mom_nnz = P.mom_nnz;
        pv.conn = P.conn.as<int>();
        pv.pos = c->pos.as<double>();
        pv.q_prev = c->q_prev.as<double>();
        pv.q_iter = w.x;
        pv.node_info = P.node_info.as<int4>();
        pv.nbr_ptr = P.nbr_ptr.as<int>();
        pv.nbr_idx = P.nbr_idx.as<int>();
        pv.pei_ptr = P.pei_ptr.as<int>();
        pv.pei = P.pei.as<uint32_t>();
        pv.dir_vals = c->dir_vals.as<double>();
        pv.vals = w.vals;
        pv.b = w.b;
        pv.vals32 = w.vals32;  // FP32 copies, when the precision mode has them
        pv.diag_pos = P.diag_pos.as<int>();
        pv.ell_nsl = P.ell_nsl;
        pv.ell_soff = P.ell_soff.as<int>();
        pv.vals_ell = w.vals_ell;
        pv.vals32_ell = w.vals32_ell;
        pv.binv = w.binv;
        pv.binv32 = w.binv32;
        pv.erec = w.erec;
        pv.esv = w.esv;
        pv.nnz = P.nnz;
        pv.rho = c->d.rho;
        pv.mu = c->d.mu;
        pv.dt = r->dt;
        for (int k = 0; k < 3; ++k) pv.g[k] = c->d.body_force[k];
        pv.problem = c->d.problem;
        pv.tau0 = c->d.tau0;
        pv.mreg = c->d.m_reg;
        if (c->d.kind == PFEM_KIND_BOUSSINESQ) {
            pv.T = c->heat.T.as<double>();
            pv.alpha = c->d.alpha_exp;
            pv.t_ref = c->d.t_ref;
        }
        pv.fst = fst_view(c, false);

        PicardView& pc = h_pic[slot];
        pc = PicardView{};
        pc.dim = dim;
        pc.N = N;
        pc.n = (int)n;
        pc.res_kind = res_kind_of(r);
        pc.dt = r->dt;
        pc.flags = P.flags.as<uint8_t>();
        pc.pos = c->pos.as<double>();
        pc.pos_save = w.pos_save;
        pc.q_prev = c->q_prev.as<double>();
        pc.x = w.x;
        pc.x_prev = w.x_prev;
        pc.A = node_mat<double>(P, dim, w.vals);
        pc.b = w.b;
        pc.partials = c->partials.as<double>();
        pc.counters = c->counters.as<unsigned>();
        pc.out = ds->out;

        KrylovView<double>& k = h_k64[slot];
        k = KrylovView<double>{};
        k.n = (int)n;
        k.x_zero = 0;
        k.nodes = N;
        k.nf = dim + 1;
        k.nblk = row_blocks(n, P.ell_nsl);
        k.A = node_mat<double>(P, dim, w.vals);
        k.binv = w.binv;
        k.b = w.b;
        k.x = w.x;
        double* v = w.kv64;
        k.r = v;
        k.rh = v + n;
        k.p = v + 2 * n;
        k.v = v + 3 * n;
        k.s = v + 4 * n;
        k.t = v + 5 * n;
        k.ph = v + 6 * n;
        k.sh = v + 7 * n;
        k.partials = c->partials.as<double>();
        k.counters = c->counters.as<unsigned>();
        k.sc = &ds->k[0];
        k.E = ell_mat<double>(P, dim, w.vals_ell);
        k.amg = c->amg.levels > 1 ? 1 : 0;
        if (k.amg) {
            if (amg_fp32(c))
                load_amg<float>(c, slot, k, h_amg32, d_amg32, node_mat<float>(P, dim, w.vals32), w.binv32);
            else {
                load_amg<double>(c, slot, k, h_amg64, d_amg64, node_mat<double>(P, dim, w.vals), w.binv);
                // K_p / K_s write the level-0 pre-smoothing (the FP64 V-cycle launches none)
                k.mg_w0 = reinterpret_cast<double*>(w.amg_w0);
                k.mg_om = std::min(amg_omega, c->amg_omega_cap);
            }
        }

        KrylovView<float>& f = h_k32[slot];
        f = KrylovView<float>{};
        RefineView& rv = h_ref[slot];
        rv = RefineView{};
        if (mixed) {
            float* u = w.kv32;
            f.n = (int)n;
            f.x_zero = 1;
            f.nodes = N;
            f.nf = dim + 1;
            f.nblk = row_blocks(n, P.ell_nsl);
            f.A = node_mat<float>(P, dim, w.vals32);
            f.E = ell_mat<float>(P, dim, w.vals32_ell);  // the inner solve's SpMVs run over the FP32 ELL copy
            f.binv = w.binv32;
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
            f.sc = &ds->k[1];
            rv.n = (int)n;
            rv.A = node_mat<double>(P, dim, w.vals);
            rv.b = w.b;
            rv.x = w.x;
            rv.r = w.kv64;
            rv.r32 = u;
            rv.d32 = u + n;
            rv.partials = c->partials.as<double>();
            rv.counters = c->counters.as<unsigned>();
            rv.out = ds->out;
        }
        upload(d_pv, h_pv, slot);
        upload(d_pic, h_pic, slot);
        upload(d_k64, h_k64, slot);
        upload(d_k32, h_k32, slot);
        upload(d_ref, h_ref, slot);
    } else {
        auto set = [&](auto& V, auto* base) {
            using S = std::remove_pointer_t<decltype(base)>;
            V = WcView<S>{};
            V.dim = dim;
            V.N = N;
            V.E = c->E;
            V.conn = c->pat.conn.as<int>();
            V.flags = c->pat.flags.as<uint8_t>();
            V.elem_ptr = c->pat.elem_ptr.as<int>();
            V.elem_list = c->pat.elem_list.as<uint32_t>();
            V.elem_nodes = c->pat.elem_nodes.as<int4>();
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
            V.hmin = acc + (3 + (size_t)dim) * N;
            V.rec = c->wrec.as<double>();
            V.erec = c->werec.as<double>();
            V.dt2 = &ds->dt2;
            V.mu = c->d.mu;
            V.K0 = c->d.K0;
            V.K0p = c->d.K0p;
            V.rho_star = c->d.rho_star;
            for (int k = 0; k < 3; ++k) V.g[k] = c->d.body_force[k];
            V.dt = r->dt;
            V.sc2 = r->sc * r->sc;
            V.meduri = c->d.meduri;
            V.fst = fst_view(c, true);
        };
        if (c->d.wc_state == PFEM_STATE_FP32) {
            set(h_wc32[slot], c->wstate.as<float>());
            upload(d_wc32, h_wc32, slot);
        } else {
            set(h_wc64[slot], c->wstate.as<double>());
            upload(d_wc64, h_wc64, slot);
        }
    }
}

void Engine::Impl::reset_kscal(int slot, int which, double rtol, int maxit, double atol) {
    KScal& k = h_reset[slot].k[which];
    std::memset(&k, 0, sizeof(KScal));
    k.rtol = rtol;
    k.atol2 = atol > 0 ? atol * atol : 0.0;
    k.maxit = maxit;
    k.req_init = 1;
    PFEM_HIP(hipMemcpyAsync(&dslot(slot)->k[which], &k, sizeof(KScal), hipMemcpyHostToDevice, s));
}

// Restart of a solve from its current iterate.  Between ticks no kernel runs, so the device
// scalars equal the last readback.
void Engine::Impl::restart_kscal(int slot, int which) {
    KScal& k = h_reset[slot].k[which];
    k = h_scal[slot].k[which];
    k.done = 0;
    k.req_init = 1;
    PFEM_HIP(hipMemcpyAsync(&dslot(slot)->k[which], &k, sizeof(KScal), hipMemcpyHostToDevice, s));
}

// Views of every level of the case's multigrid in arithmetic type T (slot * kAmgMaxLevels
// + level); A0 / binv0 are the system's values and block-Jacobi inverse in type T.
template <class T>
void Engine::Impl::load_amg(pfem_case* c, int slot, const KrylovView<double>& k, AmgView<T>* h, DevBuf& d,
                            const NodeMat<T>& A0, const T* binv0) {
    const AmgHierarchy& H = c->amg;
    const Workspace& w = c->ws;
    SlotScal* ds = dslot(slot);
    AmgView<T>* av = h + (size_t)slot * kAmgMaxLevels;
    auto arr = [](char* p) { return reinterpret_cast<T*>(p); };
    for (int l = 0; l < H.levels; ++l) {
        AmgView<T>& V = av[l];
        V = AmgView<T>{};
        V.level = l;
        V.omega = std::min(amg_omega, c->amg_omega_cap);
        V.alpha = amg_alpha;
        V.done_flag = &ds->k[0].done;
        if (l == 0) {
            V.A = A0;
            V.binv = binv0;
            if constexpr (std::is_same<T, double>::value) {
                V.E0 = ell_mat<double>(c->pat, c->dim, w.vals_ell);
                if (amg_l0_fp32(c)) V.E0f = ell_mat<float>(c->pat, c->dim, w.vals32_ell);
            } else {
                V.E0 = ell_mat<float>(c->pat, c->dim, w.vals32_ell);
            }
            V.lanes = c->lanes;
            V.mask = 1;
            V.flags = c->pat.flags.as<uint8_t>();
            V.r0[0] = k.p;
            V.r0[1] = k.s;
            V.out0[0] = k.ph;
            V.out0[1] = k.sh;
            V.w = arr(w.amg_w0);
            V.t = arr(w.amg_t0);
        } else {
            const AmgLevel& L = H.lev[l];
            NodeMat<T> A{};
            A.N = L.N;
            A.nf = c->dim + 1;
            A.mom_nnz = L.mom_nnz;
            A.info = L.node_info.as<int4>();
            A.nbr_ptr = L.nbr_ptr.as<int>();
            A.nbr_idx = L.nbr_idx.as<int>();
            A.vals = arr(w.amg_lv[l][0]);
            V.A = A;
            V.binv = arr(w.amg_lv[l][1]);
            V.lanes = L.lanes;
            V.r = arr(w.amg_lv[l][2]);
            V.w = arr(w.amg_lv[l][3]);
            V.t = arr(w.amg_lv[l][4]);
            V.z = arr(w.amg_lv[l][5]);
            V.FAP = av[l - 1].AP;  // set by level l-1 (the loop runs from the finest level)
            V.FAPf = av[l - 1].APf;
            V.gal_ptr = L.gal_ptr.as<int>();
            V.gal_src = L.gal_src.as<int2>();
            V.self_rank = L.self_rank.as<int>();
            V.vals_out = arr(w.amg_lv[l][0]);
            V.binv_out = arr(w.amg_lv[l][1]);
        }
        if (l + 1 < H.levels) {
            const AmgLevel& Lc = H.lev[l + 1];
            EllMatT<T> AP{};
            AP.N = V.A.N;
            AP.nf = c->dim + 1;
            AP.nsl = Lc.ap_nsl;
            AP.ncol = Lc.N;
            AP.soff = Lc.ap_soff.as<int>();
            AP.idx = Lc.ap_eidx.as<int>();
            AP.vals = arr(w.amg_ap[l]);
            V.ap_out = arr(w.amg_ap[l]);
            if (l == 0 && amg_l0_fp32(c)) {  // A P of level 0 in FP32 (see amg_l0_fp32)
                V.APf = EllMatT<float>{AP.N, AP.nf, AP.nsl, AP.ncol, AP.soff, AP.idx,
                                       reinterpret_cast<const float*>(w.amg_ap[l])};
                V.ap_out32 = reinterpret_cast<float*>(w.amg_ap[l]);
                AP.vals = nullptr;
                V.ap_out = nullptr;
            }
            V.AP = AP;
            V.ap_lanes = Lc.ap_lanes;
            V.ap_ptr = Lc.ap_ptr.as<int>();
            V.ap_gptr = Lc.ap_gptr.as<int>();
            V.ap_gsrc = Lc.ap_gsrc.as<int>();
            V.next_N = Lc.N;
            V.agg_ptr = Lc.agg_ptr.as<int>();
            V.agg_nodes = Lc.agg_nodes.as<int>();
            V.fine_agg = Lc.fine_agg.as<int>();
            V.next_r = arr(w.amg_lv[l + 1][2]);
            V.next_z = arr(w.amg_lv[l + 1][5]);
        } else {
            V.nc = H.coarse_n;
            V.dense = reinterpret_cast<double*>(w.amg_dense);
        }
    }
    PFEM_HIP(hipMemcpyAsync(d.as<AmgView<T>>() + (size_t)slot * kAmgMaxLevels, av, sizeof(AmgView<T>) * H.levels,
                            hipMemcpyHostToDevice, s));
}

namespace {

// Absolute cap on the linear residual of a Picard step with the absolute Ax_f residual:
// ||b - A q|| <= 0.1 minRes, so that the linear-solve error (which a direct solver keeps
// near machine precision) never decides whether ||A q - b|| <= minRes.
double step_atol(const Req* r) {
    const pfem_case_desc& d = r->c->d;
    return (r->kind == RK_STEP && d.residual == PFEM_RES_AX_F && d.min_res > 0) ? 0.1 * d.min_res : 0.0;
}

// The FP32 inner solves stopped helping (no progress, a diverging inner solve, or a
// non-finite iterate): finish with FP64 BiCGSTAB from the current iterate (from zero when
// it is not finite).  The rest of the request's solves then run in FP64 directly.
void start_fallback(Engine::Impl& I, Req* r, bool zero_x) {
    pfem_case* c = r->c;
    if (zero_x) {
        r->ops |= OP_COPY;
        r->copy |= COPY_ZERO_X;
    }
    r->fallback = true;
    r->fp64_only = true;
    r->verifying = false;
    r->restarts = 0;
    r->kit_prev = -1;
    I.reset_kscal(r->slot, 0, c->d.krylov_rtol, c->d.krylov_max_iter, step_atol(r));
    r->ops |= OP_K64_INIT | OP_K64_RUN;
    r->wait = W_K64;
}

// Operations of a solve start (FP64 BiCGSTAB or the first refinement residual).
void solve_start(Engine::Impl& I, Req* r) {
    pfem_case* c = r->c;
    r->ss = pfem_solve_stats{};
    r->verifying = false;
    r->fallback = false;
    r->restarts = 0;
    r->kit_prev = -1;
    if (c->d.precision == PFEM_PREC_MIXED_IR && !r->fp64_only) {
        r->ref = 0;
        r->stall = 0;
        r->best = std::numeric_limits<double>::infinity();
        r->ops |= OP_IRR;
        r->wait = W_IR;
    } else {
        I.reset_kscal(r->slot, 0, c->d.krylov_rtol, c->d.krylov_max_iter, step_atol(r));
        r->ops |= OP_K64_INIT | OP_K64_RUN;
        r->wait = W_K64;
    }
}

void finish(Req* r, int status, unsigned copy_ops) {
    r->status = status;
    r->final_ = true;
    if (copy_ops) {
        r->ops |= OP_COPY;
        r->copy |= copy_ops;
    }
    r->wait = W_NONE;
}

// Top of the Picard loop (PicardAlgo.cpp:47-65): `while (res > minRes)`; iterCount > maxIter
// restores the nodes and fails.
void picard_top(Engine::Impl& I, Req* r) {
    pfem_case* c = r->c;
    if (!(r->res > c->d.min_res)) {
        finish(r, PFEM_OK, COPY_COMMIT);
    } else if (r->iter > c->d.max_iter) {
        finish(r, PFEM_NOT_CONVERGED, COPY_RESTORE);
    } else {
        r->ops |= OP_COPY;
        r->copy |= COPY_SAVEIT;
        solve_start(I, r);
    }
}

void solve_done(Engine::Impl& I, Req* r, int status) {
    (void)I;
    r->ss.status = status;
    r->st.krylov_iters += r->ss.krylov_iters;
    r->st.max_relres = std::max(r->st.max_relres, r->ss.relres);
    if (r->kind == RK_SOLVE) {
        finish(r, status, 0);
        return;
    }
    // PSPG solve callback (MomContEquationPSPG.inl:278-313): failure restores the nodes
    if (status != PFEM_OK) {
        finish(r, status, COPY_RESTORE);
        return;
    }
    r->ops |= OP_UPD | OP_ASM | OP_RES;
    r->st.assemblies++;
    r->wait = W_POST;
}

// The Picard residual (MomContEquationPSPG.inl:314-372) from the sums the residual kernels
// leave in the slot: Ax_f ||A q - b||; U sqrt(du^2 / u_prev^2) over non-free nodes, DBL_MAX
// for a zero denominator; U_P the larger of that and the pressure ratio.
double picard_residual(int kind, const double* o) {
    if (kind == PFEM_RES_AX_F) return std::sqrt(o[0]);
    const double big = std::numeric_limits<double>::max();
    const double rv = (o[1] == 0) ? big : std::sqrt(o[0] / o[1]);
    const double rp = (kind == PFEM_RES_U_P) ? ((o[3] == 0) ? big : std::sqrt(o[2] / o[3])) : 0.0;
    return std::max(rv, rp);
}

// Chunk-size hint from the residual decay between two readbacks.
int estimate_remaining(Req* r, const KScal& k) {
    int est = 8;
    if (r->kit_prev >= 0 && k.iters > r->kit_prev && k.rr > 0 && r->krr_prev > k.rr && k.tol2 > 0) {
        const double rate = std::log(k.rr / r->krr_prev) / (double)(k.iters - r->kit_prev);
        const double rem = std::log(k.tol2 / k.rr) / rate;
        est = (int)std::min(1e6, std::max(1.0, std::ceil(rem)));
    }
    r->kit_prev = k.iters;
    r->krr_prev = k.rr;
    return est;
}

}  // namespace

void Engine::abort_all(int code, std::vector<Req*>& done) {
    for (Req* r : inflight_) waitq_.push_back(r);
    inflight_.clear();
    for (Req* r : waitq_) {
        r->status = code;
        r->completed = true;
        r->t_done = now_ns();
        if (r->arena_ws && p_->arena) {
            p_->arena->release(r->c->ws.off, r->c->ws.bytes);
            r->c->ws.clear();
            r->arena_ws = false;
        }
        if (r->slot >= 0) p_->free_slots.push_back(r->slot);
        r->slot = -1;
        done.push_back(r);
    }
    waitq_.clear();
}

void Engine::graph_stats(long long out[2]) const {
    out[0] = p_->graph_launches;
    out[1] = p_->graph_captures;
}

void Engine::drain(std::vector<Req*>& done) {
    while (!idle()) tick(done);
}

void Engine::tick(std::vector<Req*>& done) {
    Impl& I = *p_;
    const hipStream_t s = I.s;
    // ---- admission (FIFO): a free slot, and the step workspace of a PSPG request -------
    while (!waitq_.empty() && !I.free_slots.empty()) {
        Req* r = waitq_.front();
        pfem_case* c = r->c;
        if (pspg_like(c) && !c->ws.base) {
            const size_t wb = workspace_bytes(c);
            if (I.arena) {
                size_t off = 0;
                if (!I.arena->alloc(wb, off)) break;  // wait until memory frees up
                carve_workspace(c, I.arena->base() + off);
                c->ws.arena = I.arena;
                c->ws.off = off;
                r->arena_ws = true;
            } else {
                own_workspace(c, s);
            }
        }
        waitq_.pop_front();
        r->slot = I.free_slots.back();
        I.free_slots.pop_back();
        r->t_admit = now_ns();
        r->ops = 0;
        r->copy = 0;
        r->final_ = false;
        r->iter = 0;
        r->res = std::numeric_limits<double>::max();
        r->first_asm = true;
        r->fp64_only = false;
        r->st = pfem_step_stats{};
        I.load_views(r);
        switch (r->kind) {
            case RK_ASSEMBLE:
                r->ops = OP_ASM;
                r->final_ = true;
                break;
            case RK_SOLVE:
                solve_start(I, r);
                break;
            case RK_STEP:
                r->ops = OP_COPY | OP_ASM;
                r->copy = COPY_PREP;
                r->st.assemblies = 1;
                solve_start(I, r);
                break;
            case RK_WC:
                r->ops = (r->wc_step ? OP_WC : 0u) | (r->wc_cfl ? OP_CFL : 0u);
                r->final_ = true;
                break;
            case RK_CALLBACK:
                r->ops = r->cb_ops;
                r->copy = r->cb_copy;
                if (r->ops & OP_RES) r->wait = W_RESID;
                else r->final_ = true;
                break;
        }
        inflight_.push_back(r);
    }
    peak_inflight = std::max(peak_inflight, (int)inflight_.size());
    if (inflight_.empty()) return;
    ++ticks;

    // ---- chunk size of this tick's BiCGSTAB iterations --------------------------------
    int min_hint = 1 << 20, nrun = 0;
    for (Req* r : inflight_) {
        if (!(r->ops & (OP_K64_RUN | OP_K32_RUN))) continue;
        ++nrun;
        int hint = 8;
        if (!(r->ops & (OP_K64_INIT | OP_K32_INIT))) {
            const int w = (r->ops & OP_K64_RUN) ? 0 : 1;
            hint = estimate_remaining(r, I.h_scal[r->slot].k[w]);
        } else {
            r->kit_prev = -1;
            if (r->verifying) hint = 4;
        }
        min_hint = std::min(min_hint, hint);
    }
    const int chunk = nrun ? std::max(4, std::min(max_chunk, pow2ceil(min_hint))) : 0;

    // ---- work tables ------------------------------------------------------------------
    std::vector<BlockTask> tab[T_NTAB];
    for (Req* r : inflight_) {
        pfem_case* c = r->c;
        const int sl = r->slot;
        const unsigned op = r->ops;
        const int li = lane_index(c->lanes);
        if (op & OP_COPY) append(tab[T_COPY], sl, c->N, kElemCap, (int)r->copy);
        if (op & OP_UPD) append(tab[T_UPD], sl, c->N, kElemCap);
        if (op & OP_ASM) {
            append(tab[c->dim == 2 ? T_ASM2 : T_ASM3], sl, (long long)c->N * kAsmLanes, kElemCap);
            append(tab[T_BC], sl, c->N, kElemCap);
        }
        if (op & OP_RES) {
            if (res_kind_of(r) == PFEM_RES_AX_F) append_rows(tab[T_RES + li], sl, c->n, c->pat.ell_nsl);
            else append_rows(tab[T_RESU], sl, c->n, c->pat.ell_nsl);
        }
        if (op & OP_IRU) append_rows(tab[T_IRU], sl, c->n, c->pat.ell_nsl);
        if (op & OP_IRR) append_rows(tab[T_IRR + li], sl, c->n, c->pat.ell_nsl);
        if (op & OP_K64_RUN) append_rows(tab[T_K64 + li], sl, c->n, c->pat.ell_nsl);
        if (op & OP_K32_RUN) append_rows(tab[T_K32 + li], sl, c->n, c->pat.ell_nsl);
        if (op & OP_K64_INIT) tab[T_F64 + li].push_back(BlockTask{sl, 0, 1, 0});
        if (op & OP_K32_INIT) tab[T_F32 + li].push_back(BlockTask{sl, 0, 1, 0});
        const AmgHierarchy& H = c->amg;
        if (pspg_like(c) && H.levels > 1) {
            const int base = sl * kAmgMaxLevels;  // AmgView index of (slot, level 0)
            std::vector<BlockTask>* A = tab + T_AMG + (amg_fp32(c) ? A_N : 0);
            // the coarse operators are rebuilt at a step's first assembly; the Picard
            // re-assemblies keep them (the V-cycle stays a fixed linear preconditioner of
            // each solve) unless PFEM_AMG_REUSE=0
            if ((op & OP_ASM) && (r->first_asm || !I.amg_reuse)) {
                for (int l = 1; l < H.levels; ++l) {
                    append(A[A_G + l], base + l, (long long)H.lev[l].N * H.lev[l].lanes, kElemCap);
                    append(A[A_B + l], base + l, H.lev[l].N, kElemCap);
                }
                A[A_D].push_back(BlockTask{base + H.levels - 1, 0, 1, 0});
                for (int l = 0; l + 1 < H.levels; ++l) {
                    const int Nl = l == 0 ? c->N : H.lev[l].N;
                    append(A[A_AP + l], base + l, (long long)Nl * H.lev[l + 1].ap_lanes, kElemCap);
                }
            }
            if (op & OP_K64_RUN) {
                for (int l = 0; l + 1 < H.levels; ++l) {
                    const int Nl = l == 0 ? c->N : H.lev[l].N;
                    const int lanes = l == 0 ? c->lanes : H.lev[l].lanes;
                    append(A[A_VN + l], base + l, Nl, kElemCap);
                    if (l > 0) append(A[A_VS + l], base + l, (long long)Nl * lanes, kElemCap);
                    append(A[A_VR + l], base + l, (long long)H.lev[l + 1].N * kRestrictLanes, kElemCap);
                }
                A[A_VC].push_back(BlockTask{base + H.levels - 1, 0, 1, 0});
            }
        }
        const int wi = c->d.wc_state == PFEM_STATE_FP32 ? 1 : 0;
        if (op & OP_WC) {
            // node passes: at most kMaxCaseBlocks workgroups, each folding its CFL minimum
            // into the slot with one atomic
            append(tab[T_WCN + wi], sl, c->N, kMaxCaseBlocks);
            // node gathers over the incident elements: kWcLanes lanes per node
            append(tab[(c->dim == 2 ? T_WCE2 : T_WCE3) + wi], sl, (long long)c->N * kWcLanes[c->dim - 2], kElemCap);
            append(tab[(c->dim == 2 ? T_WEL2 : T_WEL3) + wi], sl, c->E, kElemCap);
        }
        // the step's own kernels give the CFL minimum after it; the CFL pass runs on its own
        // only without a step
        if ((op & OP_CFL) && !(op & OP_WC)) append(tab[(c->dim == 2 ? T_CFE2 : T_CFE3) + wi], sl, c->E, kElemCap);
        if (op & (OP_WC | OP_CFL)) {
            SlotScal& h = I.h_reset[sl];
            h.dt2 = 0x7fefffffffffffffULL;  // DBL_MAX
            PFEM_HIP(hipMemcpyAsync(&I.dslot(sl)->dt2, &h.dt2, sizeof(h.dt2), hipMemcpyHostToDevice, s));
        }
    }
    size_t off[T_NTAB + 1];
    off[0] = 0;
    for (int t = 0; t < T_NTAB; ++t) off[t + 1] = off[t] + tab[t].size();
    const size_t total = off[T_NTAB];
    if (total > I.tab_cap) {
        I.tab_cap = std::max<size_t>(total * 2, 4096);
        pinned_alloc(I.h_tab, I.tab_cap);
        I.d_tab.reserve(sizeof(BlockTask) * I.tab_cap);
    }
    for (int t = 0; t < T_NTAB; ++t)
        if (!tab[t].empty()) std::memcpy(I.h_tab + off[t], tab[t].data(), sizeof(BlockTask) * tab[t].size());
    if (total) PFEM_HIP(hipMemcpyAsync(I.d_tab.p, I.h_tab, sizeof(BlockTask) * total, hipMemcpyHostToDevice, s));
    auto T = [&](int t) { return I.d_tab.as<BlockTask>() + off[t]; };
    auto NT = [&](int t) { return (int)tab[t].size(); };

    // ---- launches ---------------------------------------------------------------------
    const bool timed = inflight_.size() == 1;
    Req* solo = timed ? inflight_[0] : nullptr;
    auto mark = [&](int k) {
        if (timed) PFEM_HIP(hipEventRecord(I.ev[k], s));
    };
    const PspgView* vpv = I.d_pv.as<PspgView>();
    const PicardView* vpic = I.d_pic.as<PicardView>();
    mark(0);
    if (NT(T_COPY)) launch_pic_copy(vpic, T(T_COPY), NT(T_COPY), s);
    if (NT(T_UPD)) launch_update_pos(vpic, T(T_UPD), NT(T_UPD), s);
    mark(1);
    if (NT(T_ASM2)) launch_pspg_assemble(2, vpv, T(T_ASM2), NT(T_ASM2), s);
    if (NT(T_ASM3)) launch_pspg_assemble(3, vpv, T(T_ASM3), NT(T_ASM3), s);
    if (NT(T_BC)) launch_pspg_bc(vpv, T(T_BC), NT(T_BC), s);
    // (the assembly and the BC write the sliced ELL copy and the FP32 copies themselves, and the
    // diagonal gather the block-Jacobi inverse of each node's final diagonal block)
    const AmgView<double>* vamg64 = I.d_amg64.as<AmgView<double>>();
    const AmgView<float>* vamg32 = I.d_amg32.as<AmgView<float>>();
    // multigrid setup of both hierarchy types (TA: first table of the type)
    auto amg_setup = [&](auto* v, int TA) {
        for (int l = 0; l < kAmgMaxLevels; ++l) {
            // A P of level l reads level l's final values; level l+1's Galerkin sums read it
            if (NT(TA + A_AP + l)) launch_amg_ap(v, T(TA + A_AP + l), NT(TA + A_AP + l), s);
            if (l + 1 < kAmgMaxLevels && NT(TA + A_G + l + 1)) {
                launch_amg_galerkin(v, T(TA + A_G + l + 1), NT(TA + A_G + l + 1), s);
                launch_amg_bjac(v, T(TA + A_B + l + 1), NT(TA + A_B + l + 1), s);
            }
        }
        if (NT(TA + A_D)) launch_amg_dense(v, T(TA + A_D), NT(TA + A_D), s);
    };
    amg_setup(vamg64, T_AMG);
    amg_setup(vamg32, T_AMG + A_N);
    mark(2);
    for (int li = 0; li < 3; ++li)
        if (NT(T_RES + li)) launch_res_axf(kLanes[li], vpic, T(T_RES + li), NT(T_RES + li), s);
    if (NT(T_RESU)) launch_res_u(vpic, T(T_RESU), NT(T_RESU), s);
    mark(3);
    const RefineView* vref = I.d_ref.as<RefineView>();
    if (NT(T_IRU)) launch_ir_update(vref, T(T_IRU), NT(T_IRU), s);
    for (int li = 0; li < 3; ++li)
        if (NT(T_IRR + li)) launch_ir_residual(kLanes[li], vref, T(T_IRR + li), NT(T_IRR + li), s);
    const KrylovView<double>* vk64 = I.d_k64.as<KrylovView<double>>();
    const KrylovView<float>* vk32 = I.d_k32.as<KrylovView<float>>();
    for (int li = 0; li < 3; ++li) {
        if (NT(T_F64 + li))
            launch_kry_init<double>(kLanes[li], vk64, T(T_K64 + li), NT(T_K64 + li), T(T_F64 + li), NT(T_F64 + li),
                                    I.par ^ 1, s);
        if (NT(T_F32 + li))
            launch_kry_init<float>(kLanes[li], vk32, T(T_K32 + li), NT(T_K32 + li), T(T_F32 + li), NT(T_F32 + li),
                                   I.par ^ 1, s);
    }
    // live SpMV timing: one request iterating alone
    int prof_w = -1, prof_before = 0;
    if (I.prof && nrun == 1) {
        for (Req* r : inflight_)
            if (r->ops & (OP_K64_RUN | OP_K32_RUN)) {
                prof_w = (r->ops & OP_K64_RUN) ? 0 : 1;
                const bool init = r->ops & (prof_w == 0 ? OP_K64_INIT : OP_K32_INIT);
                prof_before = init ? I.h_reset[r->slot].k[prof_w].iters : I.h_scal[r->slot].k[prof_w].iters;
            }
    }
    // multigrid V-cycle of the cases that have a hierarchy: which = 0 (ph = M^-1 p) or 1 (sh = M^-1 s).
    // Level 0's residual and every level's post-smoothing run one lane per node over ELL
    // slices (table A_VN: N_l / kBlock workgroups per case); the residual of the coarser
    // levels runs `lanes` lanes per node (A_VS).
    auto vcycle_type = [&](auto* v, int TA, int which) {
        for (int l = 0; l < kAmgMaxLevels; ++l) {
            if (!NT(TA + A_VN + l)) break;
            const int tr = TA + (l == 0 ? A_VN : A_VS) + l;
            // the pre-smoothing of levels >= 1 is part of the restriction from the finer level,
            // that of level 0 in FP64 hierarchies part of K_p / K_s (KrylovView::mg_w0)
            if (l == 0 && TA != T_AMG) launch_amg_pre(v, T(TA + A_VN + l), NT(TA + A_VN + l), which, s);
            launch_amg_res(v, T(tr), NT(tr), which, s);
            launch_amg_restrict(v, T(TA + A_VR + l), NT(TA + A_VR + l), s);
        }
        if (NT(TA + A_VC)) launch_amg_coarsest(v, T(TA + A_VC), NT(TA + A_VC), s);
        for (int l = kAmgMaxLevels - 1; l >= 0; --l)
            if (NT(TA + A_VN + l)) launch_amg_post(v, T(TA + A_VN + l), NT(TA + A_VN + l), which, s);
    };
    auto vcycle = [&](int which) {
        vcycle_type(vamg64, T_AMG, which);
        vcycle_type(vamg32, T_AMG + A_N, which);
    };
    auto run_chunk = [&]() {
        for (int it = 0; it < chunk; ++it) {
            const int p = (I.par + it) & 1;
            for (int phase = 0; phase < 5; ++phase) {
                for (int li = 0; li < 3; ++li) {
                    if (NT(T_K64 + li))
                        launch_kry_phase<double>(phase, kLanes[li], vk64, T(T_K64 + li), NT(T_K64 + li), p, s,
                                                 (phase == 1 && prof_w == 0) ? I.pev + 2 * it : nullptr);
                    if (NT(T_K32 + li))
                        launch_kry_phase<float>(phase, kLanes[li], vk32, T(T_K32 + li), NT(T_K32 + li), p, s,
                                                (phase == 1 && prof_w == 1) ? I.pev + 2 * it : nullptr);
                }
                if (phase == 0) vcycle(0);
                if (phase == 2) vcycle(1);
            }
        }
    };
    bool replayed = false;
    if (I.graphs && chunk > 0 && prof_w < 0) {
        // everything the chunk's launches capture: its length and parity, the base addresses of
        // the tables and views, and the tables' layout and contents
        std::vector<long long> key;
        key.reserve(16 + T_NTAB + 4 * total);
        key.push_back(chunk);
        key.push_back(I.par);
        for (const void* p : {I.d_tab.p, I.d_k64.p, I.d_k32.p, I.d_amg64.p, I.d_amg32.p})
            key.push_back((long long)(uintptr_t)p);
        for (int t = 0; t <= T_NTAB; ++t) key.push_back((long long)off[t]);
        for (size_t q = 0; q < total; ++q) {
            const BlockTask& b = I.h_tab[q];
            key.push_back(b.case_id);
            key.push_back(b.block);
            key.push_back(b.nblocks);
            key.push_back(b.pad);
        }
        hipGraphExec_t exec = nullptr;
        for (auto& g : I.gcache)
            if (g.key == key) exec = g.exec;
        bool again = false;
        for (const auto& k : I.seen) again = again || k == key;
        if (!exec && again) {
            hipGraph_t g = nullptr;
            PFEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            run_chunk();
            PFEM_HIP(hipStreamEndCapture(s, &g));
            const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PFEM_HIP(ie);
            if (I.gcache.size() == Impl::kGraphCache) {
                PFEM_HIP(hipGraphExecDestroy(I.gcache.front().exec));
                I.gcache.pop_front();
            }
            I.gcache.push_back(Impl::Captured{key, exec});
            ++I.graph_captures;
        }
        if (exec) {
            PFEM_HIP(hipGraphLaunch(exec, s));
            replayed = true;
            ++I.graph_launches;
        } else {
            if (I.seen.size() == Impl::kGraphSeen) I.seen.pop_front();
            I.seen.push_back(std::move(key));
        }
    }
    if (!replayed) run_chunk();
    if (nrun) I.par = (I.par + chunk) & 1;
    const WcView<double>* vw64 = I.d_wc64.as<WcView<double>>();
    const WcView<float>* vw32 = I.d_wc32.as<WcView<float>>();
    if (NT(T_WCN + 0))
        launch_wc_step<double>(vw64, T(T_WCN), NT(T_WCN), T(T_WCE2), NT(T_WCE2), T(T_WCE3), NT(T_WCE3), T(T_WEL2),
                               NT(T_WEL2), T(T_WEL3), NT(T_WEL3), s);
    if (NT(T_WCN + 1))
        launch_wc_step<float>(vw32, T(T_WCN + 1), NT(T_WCN + 1), T(T_WCE2 + 1), NT(T_WCE2 + 1), T(T_WCE3 + 1),
                              NT(T_WCE3 + 1), T(T_WEL2 + 1), NT(T_WEL2 + 1), T(T_WEL3 + 1), NT(T_WEL3 + 1), s);
    if (NT(T_CFE2) + NT(T_CFE3)) launch_wc_cfl<double>(vw64, T(T_CFE2), NT(T_CFE2), T(T_CFE3), NT(T_CFE3), s);
    if (NT(T_CFE2 + 1) + NT(T_CFE3 + 1))
        launch_wc_cfl<float>(vw32, T(T_CFE2 + 1), NT(T_CFE2 + 1), T(T_CFE3 + 1), NT(T_CFE3 + 1), s);
    mark(4);
    PFEM_HIP(hipGetLastError());
    PFEM_HIP(hipMemcpyAsync(I.h_scal, I.d_scal.p, sizeof(SlotScal) * (size_t)I.S, hipMemcpyDeviceToHost, s));
    PFEM_HIP(hipStreamSynchronize(s));

    // ---- timings ----------------------------------------------------------------------
    auto ems = [&](int a, int b) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, I.ev[a], I.ev[b]) != hipSuccess) ms = 0;
        return 1e-3 * (double)ms;
    };
    if (timed && solo->kind == RK_STEP) {
        pfem_step_stats& st = solo->st;
        if (solo->ops & OP_UPD) st.t_update += ems(0, 1);
        if (solo->ops & OP_ASM) {
            if (solo->first_asm) st.t_assemble += ems(1, 2);
            else st.t_residual += ems(1, 2);
        }
        if (solo->ops & OP_RES) st.t_residual += ems(2, 3);
        st.t_solve += ems(3, 4);
    }
    if (prof_w >= 0) {
        for (Req* r : inflight_)
            if (r->ops & (prof_w == 0 ? OP_K64_RUN : OP_K32_RUN)) {
                const int worked = std::min(chunk, I.h_scal[r->slot].k[prof_w].iters - prof_before);
                for (int it = 0; it < worked; ++it) {
                    float ms = 0;
                    if (hipEventElapsedTime(&ms, I.pev[2 * it], I.pev[2 * it + 1]) == hipSuccess) {
                        prof_ms[prof_w] += ms;
                        prof_count[prof_w]++;
                    }
                }
            }
    }

    // ---- advance every request --------------------------------------------------------
    std::vector<Req*> still;
    still.reserve(inflight_.size());
    for (Req* r : inflight_) {
        pfem_case* c = r->c;
        const unsigned ran = r->ops;
        if (ran & OP_ASM) r->first_asm = false;
        r->ops = 0;
        r->copy = 0;
        if (!r->final_) {
            const SlotScal& h = I.h_scal[r->slot];
            switch (r->wait) {
                case W_K64: {  // FP64 BiCGSTAB; convergence is accepted on a true residual only
                    const KScal& k = h.k[0];
                    r->verifying = false;
                    if (k.done == 0) {
                        r->ops = OP_K64_RUN;
                    } else if (k.done == 1 && !k.true_rr) {
                        // converged on the recurrence residual: restart from the iterate, which
                        // confirms (or not) on the true residual b - A x
                        r->verifying = true;
                        I.restart_kscal(r->slot, 0);
                        r->ops = OP_K64_INIT | OP_K64_RUN;
                    } else if (k.done == 2 && r->restarts < 8 && k.iters < k.maxit) {
                        r->restarts++;
                        I.restart_kscal(r->slot, 0);
                        r->ops = OP_K64_INIT | OP_K64_RUN;
                    } else {
                        int stt = PFEM_OK;
                        if (k.done == 2) stt = PFEM_NAN;
                        else if (k.done == 3) stt = (k.bb > 0 && std::isfinite(k.rr)) ? PFEM_NOT_CONVERGED : PFEM_NAN;
                        r->ss.krylov_iters += k.iters;
                        r->ss.restarts += r->restarts + (r->fallback ? 1 : 0);
                        r->ss.relres = k.bb > 0 ? std::sqrt(k.rr / k.bb) : 0.0;
                        solve_done(I, r, stt);
                    }
                    break;
                }
                case W_K32: {  // FP32 inner solve of one refinement sweep
                    const KScal& k = h.k[1];
                    if (k.done == 0) {
                        r->ops = OP_K32_RUN;
                    } else if (k.done == 2 && r->restarts < 8 && k.iters < k.maxit) {
                        r->restarts++;
                        I.restart_kscal(r->slot, 1);
                        r->ops = OP_K32_INIT | OP_K32_RUN;
                    } else {
                        r->ss.krylov_iters += k.iters;
                        r->ss.restarts += r->restarts;
                        // a correction is applied only when the inner solve at least halved
                        // its residual; otherwise it would corrupt the FP64 iterate
                        const bool inner_ok = std::isfinite(k.rr) && k.bb > 0 && k.rr <= 0.25 * k.bb;
                        if (inner_ok) {
                            r->ref++;
                            r->ops = OP_IRU | OP_IRR;
                            r->wait = W_IR;
                        } else {
                            start_fallback(I, r, false);
                        }
                    }
                    break;
                }
                case W_IR: {  // FP64 refinement residual
                    const double rr = h.out[0], bb = h.out[1];
                    const double relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
                    r->ss.relres = relres;
                    r->ss.refinements = r->ref;
                    const double atol = step_atol(r);
                    if (relres <= c->d.krylov_rtol && (atol <= 0 || std::sqrt(rr) <= atol)) {
                        solve_done(I, r, PFEM_OK);
                        break;
                    }
                    if (std::isfinite(relres) && relres < 0.5 * r->best) r->stall = 0;
                    else r->stall++;
                    if (std::isfinite(relres)) r->best = std::min(r->best, relres);
                    if (r->stall >= 3 || !std::isfinite(relres) || r->ref >= 40 ||
                        r->ss.krylov_iters >= c->d.krylov_max_iter) {
                        start_fallback(I, r, !std::isfinite(relres));
                    } else {
                        r->restarts = 0;
                        r->kit_prev = -1;
                        I.reset_kscal(r->slot, 1, c->d.inner_rtol, std::max(1, c->d.inner_max_iter));
                        r->ops = OP_K32_INIT | OP_K32_RUN;
                        r->wait = W_K32;
                    }
                    break;
                }
                case W_RESID: {  // pfem_residual: the value, then x_prev = x (PicardAlgo.cpp:72-73)
                    r->res = picard_residual(res_kind_of(r), h.out);
                    finish(r, PFEM_OK, COPY_SAVEIT);
                    break;
                }
                case W_POST: {  // residual callback (MomContEquationPSPG.inl:314-372)
                    const double res = picard_residual(res_kind_of(r), h.out);
                    r->res = res;
                    r->st.residual = res;
                    if (std::isnan(res)) {
                        r->st.picard_iters = r->iter + 1;
                        finish(r, PFEM_NAN, COPY_RESTORE);
                    } else {
                        r->iter++;
                        r->st.picard_iters = r->iter;
                        picard_top(I, r);
                    }
                    break;
                }
                default:
                    break;
            }
            if (r->final_ && r->ops) {  // completes after its last operations run
                still.push_back(r);
                continue;
            }
        }
        if (r->final_) {
            if (r->kind == RK_WC) {
                double dt2;
                std::memcpy(&dt2, &I.h_scal[r->slot].dt2, sizeof(dt2));
                if (r->wc_cfl) r->next_dt = std::min(std::sqrt(dt2), r->max_dt);
                // the CFL minimum of the case's current state for securityCoeff r->sc, kept
                // for a computeNextDT that follows before the state changes
                c->cfl_dt2 = dt2;
                c->cfl_sc = r->sc;
                c->cfl_valid = true;
            }
            r->t_done = now_ns();
            r->completed = true;
            if (r->arena_ws) {
                I.arena->release(c->ws.off, c->ws.bytes);
                c->ws.clear();
                r->arena_ws = false;
            }
            I.free_slots.push_back(r->slot);
            r->slot = -1;
            done.push_back(r);
            continue;
        }
        still.push_back(r);
    }
    inflight_.swap(still);
}

}  // namespace pfem
