// Host construction of the aggregation hierarchy (structure only; see amg.hpp).  It runs
// once per mesh, next to the pattern build, on the node graph copied back from the device.
#include "amg.hpp"

#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.hpp"
#include "views.hpp"

namespace pfem {

namespace {

struct Graph {
    int N = 0;
    std::vector<int> ptr, idx;  // sorted neighbour lists, self included
};

// Greedy aggregation: roots whose active neighbours are all unaggregated take those
// neighbours; the rest join the neighbouring aggregate they share most edges with; what
// is still left forms aggregates with its unaggregated neighbours.  Deterministic.
int aggregate(const Graph& G, const std::vector<char>& active, std::vector<int>& agg) {
    const int N = G.N;
    agg.assign(N, -1);
    int na = 0;
    for (int i = 0; i < N; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        bool ok = true;
        for (int m = G.ptr[i]; m < G.ptr[i + 1] && ok; ++m) {
            const int j = G.idx[m];
            if (j != i && active[j] && agg[j] >= 0) ok = false;
        }
        if (!ok) continue;
        for (int m = G.ptr[i]; m < G.ptr[i + 1]; ++m)
            if (active[G.idx[m]]) agg[G.idx[m]] = na;
        ++na;
    }
    const std::vector<int> first = agg;
    std::vector<int> cand;
    for (int i = 0; i < N; ++i) {
        if (!active[i] || first[i] >= 0) continue;
        cand.clear();
        for (int m = G.ptr[i]; m < G.ptr[i + 1]; ++m) {
            const int j = G.idx[m];
            if (j != i && first[j] >= 0) cand.push_back(first[j]);
        }
        if (cand.empty()) continue;
        std::sort(cand.begin(), cand.end());
        int best = cand[0], best_n = 0;
        for (size_t a = 0; a < cand.size();) {
            size_t b = a;
            while (b < cand.size() && cand[b] == cand[a]) ++b;
            if ((int)(b - a) > best_n) best_n = (int)(b - a), best = cand[a];
            a = b;
        }
        agg[i] = best;
    }
    for (int i = 0; i < N; ++i) {
        if (!active[i] || agg[i] >= 0) continue;
        for (int m = G.ptr[i]; m < G.ptr[i + 1]; ++m) {
            const int j = G.idx[m];
            if (active[j] && agg[j] < 0) agg[j] = na;
        }
        ++na;
    }
    return na;
}

template <class T>
void upload(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
    d.reserve(sizeof(T) * std::max<size_t>(h.size(), 1));
    if (!h.empty()) PFEM_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

int lanes_for(double avg_deg) { return avg_deg > 8 ? 16 : (avg_deg > 4 ? 8 : 4); }

}  // namespace

void build_amg(const Pattern& pat, int dim, int coarse_target, hipStream_t stream, AmgHierarchy& H) {
    H = AmgHierarchy{};
    const int nf = dim + 1;
    Graph G;
    G.N = pat.N;
    G.ptr.resize((size_t)pat.N + 1);
    G.idx.resize((size_t)pat.graph_nnz);
    std::vector<int4> info((size_t)pat.N);
    PFEM_HIP(hipMemcpyAsync(G.ptr.data(), pat.nbr_ptr.p, sizeof(int) * G.ptr.size(), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipMemcpyAsync(G.idx.data(), pat.nbr_idx.p, sizeof(int) * G.idx.size(), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipMemcpyAsync(info.data(), pat.node_info.p, sizeof(int4) * info.size(), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipStreamSynchronize(stream));
    // level 0: free nodes (identity rows, no element) stay out of the hierarchy
    std::vector<char> active((size_t)G.N);
    for (int i = 0; i < G.N; ++i) active[i] = (info[i].w & NODE_FREE) == 0;

    std::vector<AmgLevel> levels(1);
    std::vector<int> agg;
    while ((int)levels.size() < kAmgMaxLevels) {
        int n_active = 0;
        for (char a : active) n_active += a;
        if (n_active <= coarse_target) break;
        const int na = aggregate(G, active, agg);
        if (na == 0 || na > (4 * n_active) / 5) break;
        // coarse graph
        std::vector<unsigned long long> keys;
        keys.reserve(G.idx.size());
        for (int i = 0; i < G.N; ++i) {
            if (!active[i]) continue;
            for (int m = G.ptr[i]; m < G.ptr[i + 1]; ++m) {
                const int j = G.idx[m];
                if (active[j]) keys.push_back((unsigned long long)agg[i] * (unsigned long long)na + agg[j]);
            }
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        Graph C;
        C.N = na;
        C.ptr.assign((size_t)na + 1, 0);
        C.idx.resize(keys.size());
        for (size_t m = 0; m < keys.size(); ++m) {
            C.ptr[keys[m] / na + 1]++;
            C.idx[m] = (int)(keys[m] % na);
        }
        for (int I = 0; I < na; ++I) C.ptr[I + 1] += C.ptr[I];
        // A P structure: for each finer node i, the distinct aggregates of its active
        // neighbours (sorted) and, per (i, J), the ranks k of those neighbours in i's list
        std::vector<int> ap_ptr((size_t)G.N + 1, 0), ap_idx, ap_gptr(1, 0), ap_gsrc;
        {
            std::vector<std::pair<int, int>> nb;  // (aggregate, rank)
            int apmax = 0;
            for (int i = 0; i < G.N; ++i) {
                nb.clear();
                if (active[i])
                    for (int m = G.ptr[i]; m < G.ptr[i + 1]; ++m) {
                        const int j = G.idx[m];
                        if (active[j] && agg[j] >= 0) nb.emplace_back(agg[j], m - G.ptr[i]);
                    }
                std::sort(nb.begin(), nb.end());
                for (size_t a = 0; a < nb.size();) {
                    size_t b = a;
                    while (b < nb.size() && nb[b].first == nb[a].first) ap_gsrc.push_back(nb[b++].second);
                    ap_idx.push_back(nb[a].first);
                    ap_gptr.push_back((int)ap_gsrc.size());
                    a = b;
                }
                ap_ptr[i + 1] = (int)ap_idx.size();
                apmax = std::max(apmax, ap_ptr[i + 1] - ap_ptr[i]);
            }
            PFEM_REQUIRE(apmax <= 255, E_UNSUPPORTED, "a node's neighbours span more than 255 aggregates");
        }
        // Galerkin gather lists: the A P entries (i, m) of each coarse entry (I, J), I = agg[i]
        // and J the aggregate of i's rank-m entry, in finer node order
        std::vector<int> gcount(C.idx.size() + 1, 0);
        std::vector<int> centry(ap_idx.size(), -1);
        for (int i = 0; i < G.N; ++i) {
            if (!active[i]) continue;
            const int I = agg[i];
            const int* b = C.idx.data() + C.ptr[I];
            const int* e = C.idx.data() + C.ptr[I + 1];
            for (int m = ap_ptr[i]; m < ap_ptr[i + 1]; ++m) {
                const int* f = std::lower_bound(b, e, ap_idx[m]);
                PFEM_REQUIRE(f != e && *f == ap_idx[m], E_STATE, "A P entry outside the coarse graph");
                const int mc = (int)(f - C.idx.data());
                centry[m] = mc;
                gcount[mc + 1]++;
            }
        }
        for (size_t m = 0; m + 1 < gcount.size(); ++m) gcount[m + 1] += gcount[m];
        std::vector<int2> gsrc((size_t)gcount.back());
        std::vector<int> cursor(gcount.begin(), gcount.end() - 1);
        for (int i = 0; i < G.N; ++i)
            for (int m = ap_ptr[i]; m < ap_ptr[i + 1]; ++m)
                if (centry[m] >= 0) gsrc[cursor[centry[m]]++] = int2{i, m - ap_ptr[i]};
        // aggregate members
        std::vector<int> aptr((size_t)na + 1, 0), anodes;
        for (int i = 0; i < G.N; ++i)
            if (active[i]) aptr[agg[i] + 1]++;
        for (int I = 0; I < na; ++I) aptr[I + 1] += aptr[I];
        anodes.resize((size_t)aptr[na]);
        {
            std::vector<int> cur(aptr.begin(), aptr.end() - 1);
            for (int i = 0; i < G.N; ++i)
                if (active[i]) anodes[cur[agg[i]]++] = i;
        }
        // node descriptors of the coarse level (every row full)
        std::vector<int4> cinfo((size_t)na);
        std::vector<int> self((size_t)na);
        int maxdeg = 0;
        for (int I = 0; I < na; ++I) {
            const int deg = C.ptr[I + 1] - C.ptr[I];
            maxdeg = std::max(maxdeg, deg);
            cinfo[I] = int4{nf * C.ptr[I], nf * C.ptr[I], deg, 0};
            const int* b = C.idx.data() + C.ptr[I];
            self[I] = (int)(std::lower_bound(b, b + deg, I) - b);
        }
        PFEM_REQUIRE(maxdeg <= 255, E_UNSUPPORTED, "aggregation produced a coarse node degree above 255");
        AmgLevel L;
        L.N = na;
        L.graph_nnz = (int)C.idx.size();
        L.max_deg = maxdeg;
        L.lanes = lanes_for((double)L.graph_nnz / na);
        L.mom_nnz = nf * L.graph_nnz;
        L.nnz = (long long)nf * nf * L.graph_nnz;
        upload(L.nbr_ptr, C.ptr, stream);
        upload(L.nbr_idx, C.idx, stream);
        upload(L.node_info, cinfo, stream);
        upload(L.self_rank, self, stream);
        upload(L.agg_ptr, aptr, stream);
        upload(L.agg_nodes, anodes, stream);
        upload(L.fine_agg, agg, stream);
        upload(L.gal_ptr, gcount, stream);
        upload(L.gal_src, gsrc, stream);
        L.ap_nnz = (int)ap_idx.size();
        L.ap_lanes = lanes_for(G.N ? (double)L.ap_nnz / G.N : 0.0);
        {  // sliced ELL structure of A P
            const int nsl = (G.N + kEllSlice - 1) / kEllSlice;
            std::vector<int> soff((size_t)nsl + 1, 0);
            for (int c = 0; c < nsl; ++c) {
                int w = 0;
                for (int r = 0; r < kEllSlice; ++r) {
                    const int i = c * kEllSlice + r;
                    if (i < G.N) w = std::max(w, ap_ptr[i + 1] - ap_ptr[i]);
                }
                soff[c + 1] = soff[c] + w;
            }
            std::vector<int> eidx((size_t)soff[nsl] * kEllSlice, 0);
            for (int c = 0; c < nsl; ++c)
                for (int r = 0; r < kEllSlice; ++r) {
                    const int i = c * kEllSlice + r;
                    if (i >= G.N) continue;
                    for (int m = 0; m < ap_ptr[i + 1] - ap_ptr[i]; ++m)
                        eidx[((size_t)soff[c] + m) * kEllSlice + r] = ap_idx[(size_t)ap_ptr[i] + m];
                }
            L.ap_nsl = nsl;
            L.ap_ranks = soff[nsl];
            upload(L.ap_soff, soff, stream);
            upload(L.ap_eidx, eidx, stream);
        }
        upload(L.ap_ptr, ap_ptr, stream);
        upload(L.ap_gptr, ap_gptr, stream);
        upload(L.ap_gsrc, ap_gsrc, stream);
        PFEM_HIP(hipStreamSynchronize(stream));
        levels.push_back(std::move(L));
        G = std::move(C);
        active.assign((size_t)na, 1);
    }
    if (levels.size() < 2) return;  // nothing to coarsen: block Jacobi alone
    const int ncoarse = levels.back().N * nf;
    if (ncoarse > kAmgMaxDense) return;  // the coarsest level must fit the dense inverse in LDS
    H.levels = (int)levels.size();
    H.lev = std::move(levels);
    H.coarse_n = ncoarse;
}

}  // namespace pfem
