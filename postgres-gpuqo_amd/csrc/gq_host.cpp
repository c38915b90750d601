/*
 * gq_host.cpp -- C ABI of the optimizer core and the host-side drivers.
 *
 * gq_plan follows __gpuqo_run (src/backend/optimizer/gpuqo/
 * gpuqo_main_internal.cu:81-147): convert the input, relabel relations in
 * BFS order from relation 0, run the DP (or IDP1 / IDP2 / UnionDP), map the
 * tree back.  The DPs run on the GPU (gq_dp.hip); the drivers, the Remapper
 * and GOO run here.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <queue>
#include <unordered_map>

#include "gq_drivers.h"
#include "gq_host.h"
#include "gq_wide.h"

namespace gq {

/* ------------------------------------------------------------ errors */
static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

static std::atomic<int> g_interrupt(0);
bool interrupted() { return g_interrupt.load() != 0; }

/* ------------------------------------------------------------ cost view */

template <int W>
void HostView<W>::build(const HQuery<W> &q) {
    const int n = q.n, n_ec = (int)q.ec_relids.size();
    const int ecw = std::max(1, (n_ec + 63) / 64);
    rel_rows = q.rows;
    rel_fksel.resize(n);
    for (int i = 0; i < n; i++)   /* fkselc_of_ec: double division narrowed to float (row_estimation.cuh:39) */
        rel_fksel[i] = (float)(1.0 / (double)fmaxf(1.0f, q.tuples[i]));
    inc.assign((size_t)n * ecw, 0);
    ec_moff.resize(n_ec);
    ec_soff.resize(n_ec);
    ec_hasfk.assign(n_ec, 0);
    mem_rel.clear();
    int moff = 0, soff = 0;
    for (int e = 0; e < n_ec; e++) {
        ec_moff[e] = moff;
        ec_soff[e] = soff;
        Set<W> t = q.ec_relids[e];
        int s = popc(t);
        while (!empty(t)) {
            int r = ctz(t);
            t.unset(r);
            inc[(size_t)r * ecw + (e >> 6)] |= 1ull << (e & 63);
            mem_rel.push_back(r);
        }
        for (int m = moff; m < moff + s; m++)
            if (!empty(q.fks[m])) ec_hasfk[e] = 1;
        moff += s;
        soff += s * (s - 1) / 2;
    }
    const int nm = moff;
    mem_nd.assign(nm, 0.0f);
    mem_inv_nd.assign(nm, 0.0f);
    mem_skew.assign(nm, 1.0f);
    mem_mcv.assign(nm, 0.0f);
    mem_flags.assign(nm, 0);
    mem_idx_rows.assign(nm, 0.0f);
    mem_idx_startup.assign(nm, 0.0f);
    mem_idx_total.assign(nm, 0.0f);
    for (int m = 0; m < nm; m++) {
        const gq_var_info &var = q.vars[m];
        const int r = mem_rel[m];
        /* __estimate_hash_bucketsize up to the nbuckets test (cost_postgres.cuh:711-742) */
        float nd = var.stadistinct;
        if (nd < 0.0f) nd = -nd * q.tuples[r];
        if (nd <= 0.0f) {
            mem_flags[m] |= GQ_MF_BUCKET_DEFAULT;
        } else {
            float avgfreq = (1.0f - var.stanullfrac) / nd;
            if (q.tuples[r] > 0.0f) {
                nd *= q.rows[r] / q.tuples[r];
                nd = clamp_row_est(nd);
            }
            mem_nd[m] = nd;
            mem_inv_nd[m] = 1.0f / nd;                              /* :751-754 */
            if (avgfreq > 0.0f && var.mcvfreq > avgfreq)            /* :766-767 */
                mem_skew[m] = var.mcvfreq / avgfreq;
        }
        mem_mcv[m] = var.mcvfreq;
        if (var.index_available) mem_flags[m] |= GQ_MF_INDEX;
        if (var.index_unique) mem_flags[m] |= GQ_MF_UNIQUE;
        mem_idx_rows[m] = var.index_rows;
        mem_idx_startup[m] = var.index_cost.startup;
        mem_idx_total[m] = var.index_cost.total;
    }
    qualtab.assign(n_ec + 1, 0.0f);
    float acc = 0.0f;                /* cost_qual_eval: repeated addition (cost_postgres.cuh:801-812) */
    for (int k = 1; k <= n_ec; k++) {
        acc += q.p.cpu_operator_cost;
        qualtab[k] = acc;
    }
    v.n_rels = n;
    v.n_ec = n_ec;
    v.ecw = ecw;
    v.p = q.p;
    v.rel_rows = rel_rows.data();
    v.rel_fksel = rel_fksel.data();
    Set<W> comp;
    for (int i = 0; i < n; i++)
        if (q.composite[i]) comp.set(i);
    v.composite = comp;
    v.inc = inc.data();
    v.ec_relids = q.ec_relids.data();
    v.ec_moff = ec_moff.data();
    v.ec_soff = ec_soff.data();
    v.ec_hasfk = ec_hasfk.data();
    v.mem_fks = q.fks.data();
    v.mem_nd = mem_nd.data();
    v.mem_inv_nd = mem_inv_nd.data();
    v.mem_skew = mem_skew.data();
    v.mem_mcv = mem_mcv.data();
    v.mem_flags = mem_flags.data();
    v.mem_idx_rows = mem_idx_rows.data();
    v.mem_idx_startup = mem_idx_startup.data();
    v.mem_idx_total = mem_idx_total.data();
    v.sels = q.sels.data();
    v.qualtab = qualtab.data();
    /* two-member classes: the selectivity of each orientation, from the
     * same estimate_ec_selectivity code (FK test included) */
    v.pairs = 0;
    v.pair_rec = nullptr;
    bool pairs = n_ec > 0;
    for (int e = 0; e < n_ec; e++)
        if (popc(q.ec_relids[e]) != 2) pairs = false;
    if (pairs) {
        pair_rec.resize((size_t)3 * n_ec);
        for (int e = 0; e < n_ec; e++) {
            Set<W> lo = lowbit(q.ec_relids[e]), hi = q.ec_relids[e] - lo;
            PairRec &c = pair_rec[3 * (size_t)e];
            c.x = ec_selectivity(v, e, lo, hi);
            c.y = ec_selectivity(v, e, hi, lo);
            c.z = (float)ctz(lo);
            c.w = 0.0f;
            for (int k = 0; k < 2; k++) {
                const int m = ec_moff[e] + k;
                PairRec &r = pair_rec[3 * (size_t)e + 1 + k];
                if (mem_flags[m] & GQ_MF_BUCKET_DEFAULT) {
                    r.x = -INFINITY;
                    r.y = 0.1f;
                    r.z = 1.0f;
                } else {
                    r.x = mem_nd[m];
                    r.y = mem_inv_nd[m];
                    r.z = mem_skew[m];
                }
                r.w = mem_mcv[m];
            }
        }
        v.pairs = 1;
        v.pair_rec = pair_rec.data();
    }
}

template struct HostView<1>;
template struct HostView<4>;

/* ------------------------------------------------------------ device blob */

/* The device's view of a blob (make_view in gq_dp.hip), built on the host
 * over the same words so that host-side precomputation runs the same code
 * on the same data. */
static View<uint32_t> host_view(const BlobHdr &h, const uint32_t *s) {
    View<uint32_t> v;
    v.n_rels = h.n_rels;
    v.n_ec = h.n_ec;
    v.ecw = h.ecw;
    v.p = h.p;
    v.rel_rows = (const float *)(s + h.o_rel_rows);
    v.rel_fksel = (const float *)(s + h.o_rel_fksel);
    v.composite = h.composite;
    v.inc = (const uint64_t *)(s + h.o_inc);
    v.ec_relids = s + h.o_ec_relids;
    v.ec_moff = (const int *)(s + h.o_ec_moff);
    v.ec_soff = (const int *)(s + h.o_ec_soff);
    v.ec_hasfk = (const uint8_t *)(s + h.o_ec_hasfk);
    v.mem_fks = s + h.o_mem_fks;
    v.mem_nd = (const float *)(s + h.o_mem_nd);
    v.mem_inv_nd = (const float *)(s + h.o_mem_inv_nd);
    v.mem_skew = (const float *)(s + h.o_mem_skew);
    v.mem_mcv = (const float *)(s + h.o_mem_mcv);
    v.mem_flags = (const uint8_t *)(s + h.o_mem_flags);
    v.mem_idx_rows = (const float *)(s + h.o_mem_idx_rows);
    v.mem_idx_startup = (const float *)(s + h.o_mem_idx_startup);
    v.mem_idx_total = (const float *)(s + h.o_mem_idx_total);
    v.sels = (const float *)(s + h.o_sels);
    v.qualtab = (const float *)(s + h.o_qualtab);
    v.pairs = h.pairs;
    v.pair_rec = (const PairRec *)(s + h.o_pair_rec);
    return v;
}

/* Per-edge split table of a tree query (BlobHdr::tree_fast): for the split
 * at edge (p, c), the crossing classes are those on that edge (o_ecm), so
 * estimate_join_selectivity reduces to the product, in ascending class
 * order, of ec_selectivity on the one-member sides, and the inner member
 * slots to those of c (o = 0, c on the inner side) or p (o = 1). */
static void fill_tree_edges(const BlobHdr &h, std::vector<uint32_t> &b) {
    const View<uint32_t> q = host_view(h, b.data());
    for (int c = 1; c < h.n_rels; c++) {
        /* the parent: the one lower-numbered neighbour (BlobHdr::is_tree) */
        const int p = __builtin_ctz(b[h.o_adj + c] & ((1u << c) - 1u));
        const uint64_t *em = (const uint64_t *)(b.data() + h.o_ecm) + (size_t)c * h.ecw;
        const uint32_t sc = 1u << c, sp = 1u << p;
        float sel0 = 1.0f, sel1 = 1.0f;
        uint16_t mi0[4] = {0, 0, 0, 0}, mi1[4] = {0, 0, 0, 0};
        int k = 0, nq = 0;
        for (int w = 0; w < h.ecw; w++) {
            uint64_t m = em[w];
            while (m) {
                const int e = w * 64 + __builtin_ctzll(m);
                m &= m - 1;
                sel0 *= ec_selectivity(q, e, sp, sc);
                sel1 *= ec_selectivity(q, e, sc, sp);
                const uint32_t ec = q.ec_relids[e];
                mi0[k] = (uint16_t)(q.ec_moff[e] + rank_below(ec, sc));
                mi1[k] = (uint16_t)(q.ec_moff[e] + rank_below(ec, sp));
                k++;
                nq++;
            }
        }
        uint32_t *t = b.data() + h.o_tedge + 8 * c;
        memcpy(&t[0], &sel0, 4);
        memcpy(&t[1], &sel1, 4);
        t[2] = (uint32_t)nq | ((uint32_t)k << 8);
        t[3] = 0;
        memcpy(&t[4], mi0, 8);
        memcpy(&t[6], mi1, 8);
    }
}

/* Plain DPsub on a sparse cyclic graph: a split (L, S - L) with both sides
 * connected cuts edges of one biconnected block of S only (two cut edges in
 * different blocks would close a cycle through both), so the MPDP block
 * enumeration visits exactly DPsub's valid splits; ranked by DPsub's
 * j = pext(L, S) they give DPsub's plan while the 2^k - 2 subsets of a set
 * are never tested.  Taken when the blocks of the whole graph (a set's
 * blocks never exceed them) enumerate at most 2^n / 64 splits -- m(m - 1)
 * arcs of a simple cycle of m relations, 2^m - 2 subsets of any other block
 * -- and no block but a simple cycle has more than 16 relations. */
static bool dpsub_by_blocks(int n, const std::vector<Set<1>> &adjs) {
    if (n < 12 || n > 31) return false;
    uint32_t adj[32];
    for (int i = 0; i < n; i++) adj[i] = (uint32_t)adjs[i].w[0];
    int disc[32], low[32], timer = 0;
    for (int i = 0; i < n; i++) disc[i] = -1;
    std::vector<std::pair<int, int>> st;
    uint64_t items = 0;
    bool ok = true;
    std::function<void(int, int)> dfs = [&](int u, int p) {
        disc[u] = low[u] = timer++;
        for (uint32_t r = adj[u]; r; r &= r - 1u) {
            const int v = __builtin_ctz(r);
            if (disc[v] < 0) {
                st.emplace_back(u, v);
                dfs(v, u);
                low[u] = std::min(low[u], low[v]);
                if (low[v] >= disc[u]) {                /* u separates the block holding edge (u, v) */
                    uint32_t vs = 0;
                    int edges = 0;
                    while (true) {
                        const std::pair<int, int> e = st.back();
                        st.pop_back();
                        vs |= (1u << e.first) | (1u << e.second);
                        edges++;
                        if (e.first == u && e.second == v) break;
                    }
                    const int m = __builtin_popcount(vs);
                    const bool cycle = m >= 3 && edges == m;
                    if (!cycle && m > 16) ok = false;
                    items += cycle ? (uint64_t)m * (m - 1) : (1ull << m) - 2ull;
                }
            } else if (v != p && disc[v] < disc[u]) {
                st.emplace_back(u, v);
                low[u] = std::min(low[u], disc[v]);
            }
        }
    };
    dfs(0, -1);
    for (int i = 0; i < n; i++) ok = ok && disc[i] >= 0;     /* connected */
    return ok && items * 64 <= (1ull << n);
}

int build_dp_input(const HQuery<1> &q, DPInput &in) {
    const int n = q.n;
    if (n > 32) {
        set_error("GPU DP supports at most 32 relations (got %d)", n);
        return GQ_ERR_UNSUPPORTED;
    }
    HostView<1> hv;
    hv.build(q);
    const int n_ec = hv.v.n_ec, ecw = hv.v.ecw, nm = (int)q.vars.size();
    if (ecw > GQ_MAX_ECW) {
        set_error("GPU DP supports at most %d equivalence classes (got %d)", 64 * GQ_MAX_ECW, n_ec);
        return GQ_ERR_UNSUPPORTED;
    }
    BlobHdr &h = in.hdr;
    std::vector<uint32_t> &b = in.blob;
    b.clear();
    auto take = [&](int words) {
        int o = (int)b.size();
        b.resize(b.size() + words, 0u);
        return o;
    };
    auto align2 = [&]() { if (b.size() & 1) b.push_back(0u); };
    auto put_f = [&](int o, int i, float x) { memcpy(&b[o + i], &x, 4); };
    h.n_rels = n;
    h.n_ec = n_ec;
    h.n_mem = nm;
    h.n_sels = (int)q.sels.size();
    h.ecw = ecw;
    h.p = q.p;
    h.composite = (uint32_t)hv.v.composite.w[0];
    h.o_rel_rows = take(32);
    h.o_rel_fksel = take(32);
    h.o_adj = take(32);
    for (int i = 0; i < n; i++) {
        put_f(h.o_rel_rows, i, hv.rel_rows[i]);
        put_f(h.o_rel_fksel, i, hv.rel_fksel[i]);
        b[h.o_adj + i] = (uint32_t)q.adj[i].w[0];
    }
    align2();
    h.o_inc = take(2 * 32 * ecw);
    for (int i = 0; i < n; i++)
        for (int w = 0; w < ecw; w++) {
            uint64_t x = hv.inc[(size_t)i * ecw + w];
            b[h.o_inc + 2 * (i * ecw + w)] = (uint32_t)x;
            b[h.o_inc + 2 * (i * ecw + w) + 1] = (uint32_t)(x >> 32);
        }
    h.o_ec_relids = take(std::max(1, n_ec));
    h.o_ec_moff = take(std::max(1, n_ec));
    h.o_ec_soff = take(std::max(1, n_ec));
    h.o_ec_hasfk = take(std::max(1, (n_ec + 3) / 4));
    for (int e = 0; e < n_ec; e++) {
        b[h.o_ec_relids + e] = (uint32_t)q.ec_relids[e].w[0];
        b[h.o_ec_moff + e] = (uint32_t)hv.ec_moff[e];
        b[h.o_ec_soff + e] = (uint32_t)hv.ec_soff[e];
        ((uint8_t *)&b[h.o_ec_hasfk])[e] = hv.ec_hasfk[e];
    }
    h.o_mem_fks = take(std::max(1, nm));
    h.o_mem_nd = take(std::max(1, nm));
    h.o_mem_inv_nd = take(std::max(1, nm));
    h.o_mem_skew = take(std::max(1, nm));
    h.o_mem_mcv = take(std::max(1, nm));
    h.o_mem_flags = take(std::max(1, (nm + 3) / 4));
    h.o_mem_idx_rows = take(std::max(1, nm));
    h.o_mem_idx_startup = take(std::max(1, nm));
    h.o_mem_idx_total = take(std::max(1, nm));
    for (int m = 0; m < nm; m++) {
        b[h.o_mem_fks + m] = (uint32_t)q.fks[m].w[0];
        put_f(h.o_mem_nd, m, hv.mem_nd[m]);
        put_f(h.o_mem_inv_nd, m, hv.mem_inv_nd[m]);
        put_f(h.o_mem_skew, m, hv.mem_skew[m]);
        put_f(h.o_mem_mcv, m, hv.mem_mcv[m]);
        ((uint8_t *)&b[h.o_mem_flags])[m] = hv.mem_flags[m];
        put_f(h.o_mem_idx_rows, m, hv.mem_idx_rows[m]);
        put_f(h.o_mem_idx_startup, m, hv.mem_idx_startup[m]);
        put_f(h.o_mem_idx_total, m, hv.mem_idx_total[m]);
    }
    h.o_sels = take(std::max<int>(1, (int)q.sels.size()));
    for (size_t i = 0; i < q.sels.size(); i++) put_f(h.o_sels, (int)i, q.sels[i]);
    h.o_qualtab = take(n_ec + 1);
    for (int i = 0; i <= n_ec; i++) put_f(h.o_qualtab, i, hv.qualtab[i]);
    h.pairs = hv.v.pairs;
    while (b.size() & 3) b.push_back(0u);       /* 16-byte records: 4-word aligned in the staged blob */
    h.o_pair_rec = take(h.pairs ? 12 * n_ec : 4);
    if (h.pairs) memcpy(&b[h.o_pair_rec], hv.pair_rec.data(), sizeof(PairRec) * 3 * (size_t)n_ec);
    in.tree_counts.clear();
    in.memo = GQ_MEMO_DENSE;
    h.o_desc = take(32);
    h.o_lglob = take(32);
    h.tree_ecm = 0;
    h.o_ecm = 0;
    h.is_tree = n >= 2 ? 1 : 0;
    for (int v = 1; v < n; v++)
        if (popc((uint32_t)q.adj[v].w[0] & ((1u << v) - 1u)) != 1) h.is_tree = 0;
    if (h.is_tree) {
        std::vector<int> parent(n, -1);
        std::vector<uint32_t> desc(n, 0), lg(n, 0);
        for (int v = 1; v < n; v++) parent[v] = ctz((uint32_t)q.adj[v].w[0] & ((1u << v) - 1u));
        for (int v = n - 1; v >= 0; v--) {
            desc[v] |= 1u << v;
            if (v > 0) desc[parent[v]] |= desc[v];
        }
        for (int v = 1; v < n; v++) {          /* parents precede children */
            uint32_t l = lg[parent[v]];
            for (int s2 = 1; s2 < v; s2++)
                if (parent[s2] == parent[v]) l |= desc[s2];
            lg[v] = l;
        }
        for (int v = 0; v < n; v++) {
            b[h.o_desc + v] = desc[v];
            b[h.o_lglob + v] = lg[v];
        }
        /* crossing classes per tree edge, when every class is on one edge */
        std::vector<uint64_t> ecm((size_t)n * ecw, 0ull);
        bool ok = true;
        for (int e = 0; e < n_ec && ok; e++) {
            const uint32_t rel = (uint32_t)q.ec_relids[e].w[0];
            const int pc = popc(rel);
            if (pc < 2) continue;
            if (pc != 2) { ok = false; break; }
            const int u = ctz(rel), v = 31 - __builtin_clz(rel);   /* u < v */
            if (parent[v] != u) { ok = false; break; }
            ecm[(size_t)v * ecw + e / 64] |= 1ull << (e % 64);
        }
        /* connected sets per size: P_v(x) = x * prod_{c child of v} (1 + P_c(x)),
         * summed over v (each subtree counted at its top vertex) */
        {
            std::vector<std::vector<uint64_t>> P(n);
            std::vector<uint64_t> tot(n + 1, 0ull);
            for (int v = n - 1; v >= 0; v--) {
                std::vector<uint64_t> poly(n + 1, 0ull);
                poly[1] = 1;
                for (int c = v + 1; c < n; c++) {
                    if (parent[c] != v) continue;
                    std::vector<uint64_t> q(P[c]);
                    q[0] += 1;
                    std::vector<uint64_t> r(n + 1, 0ull);
                    for (int a = 0; a <= n; a++) {
                        if (!poly[a]) continue;
                        for (int b = 0; a + b <= n; b++) r[a + b] += poly[a] * q[b];
                    }
                    poly.swap(r);
                }
                P[v] = poly;
                for (int k = 0; k <= n; k++) tot[k] += poly[k];
            }
            in.tree_counts = tot;
        }
        if (ok) {
            align2();
            h.o_ecm = take(2 * n * ecw);
            for (int v = 0; v < n; v++)
                for (int w = 0; w < ecw; w++) {
                    const uint64_t x = ecm[(size_t)v * ecw + w];
                    b[h.o_ecm + 2 * (v * ecw + w)] = (uint32_t)x;
                    b[h.o_ecm + 2 * (v * ecw + w) + 1] = (uint32_t)(x >> 32);
                }
            h.tree_ecm = 1;
        }
    }
    /* per-edge split table of tree graphs (see BlobHdr::tree_fast) */
    h.tree_fast = 0;
    h.o_tedge = 0;
    h.o_rank = 0;
    if (h.is_tree && h.tree_ecm) {
        bool fits = true;
        for (int v = 1; v < n && fits; v++) {
            int cnt = 0;
            for (int w = 0; w < ecw; w++)
                cnt += __builtin_popcountll(((const uint64_t *)&b[h.o_ecm])[v * ecw + w]);
            fits = cnt <= 4;
        }
        if (fits) {
            h.o_tedge = take(8 * n);
            h.tree_fast = 1;
        }
    }
    h.o_base = take(4 * n);
    in.base.resize(n);
    for (int i = 0; i < n; i++) {
        float st = q.startup[i], tot = q.total[i];
        base_cost(q.p, q.tuples[i], q.composite[i] != 0, &st, &tot);   /* cost_baserel */
        put_f(h.o_base, 4 * i, q.rows[i]);
        put_f(h.o_base, 4 * i + 1, st);
        put_f(h.o_base, 4 * i + 2, tot);
        b[h.o_base + 4 * i + 3] = (uint32_t)q.width[i];
        in.base[i] = Rel{q.rows[i], st, tot, q.width[i]};
    }
    while (b.size() & 3) b.push_back(0u);
    h.words = (int)b.size();
    if (h.tree_fast) fill_tree_edges(h, b);
    in.dpsub_blocks = !h.is_tree && dpsub_by_blocks(n, q.adj);
    /* Level-1 list of cpu_dpsize: iteration order of a libstdc++
     * unordered_map holding the base relations, inserted in index order, keyed
     * by the reference's relation ids (relation i -> bit i+1), whose hash is
     * the identity (gpuqo_cpu_dpsize.cu:43-47, gpuqo_cpu_sequential.cu:73-85). */
    std::unordered_map<uint64_t, int> m;
    for (int i = 0; i < n; i++) m.insert(std::make_pair(1ull << (i + 1), i));
    in.level1.clear();
    for (auto &kv : m) in.level1.push_back(kv.second);
    return GQ_OK;
}

/* ------------------------------------------------------------ conversion */

static Params make_params(const gq_config *c) {     /* setParams, gpuqo_planner_info.cu:113-131 */
    Params p;
    p.seq_page_cost = (float)c->seq_page_cost;
    p.cpu_tuple_cost = (float)c->cpu_tuple_cost;
    p.cpu_operator_cost = (float)c->cpu_operator_cost;
    p.disable_cost = (float)c->disable_cost;
    p.work_mem = c->work_mem;
    p.enable_nestloop = c->enable_nestloop;
    p.enable_hashjoin = c->enable_hashjoin;
    p.cost_fn = c->cost_function;
    return p;
}

template <int W>
static Set<W> set_from_words(const uint64_t *w, int nw) {
    Set<W> s;
    for (int i = 0; i < nw && i < W; i++) s.w[i] = w[i];
    return s;
}

int validate_config(const gq_config *c) {
    if (c->cost_function != GQ_COST_POSTGRES && c->cost_function != GQ_COST_SIMPLE &&
        c->cost_function != GQ_COST_COUT) {
        set_error("cost_function %d is not GQ_COST_POSTGRES, GQ_COST_SIMPLE or GQ_COST_COUT", c->cost_function);
        return GQ_ERR_INVALID;
    }
    return GQ_OK;
}

int validate(const gq_query *q) {
    if (!q || q->n_rels < 1 || q->n_rels > GQ_MAX_RELS) {
        set_error("n_rels must be in 1..%d", GQ_MAX_RELS);
        return GQ_ERR_INVALID;
    }
    if (q->n_words != (q->n_rels + 63) / 64) {
        set_error("n_words must be ceil(n_rels/64)");
        return GQ_ERR_INVALID;
    }
    if (!q->base_rels || !q->edge_table || (q->n_eq_classes > 0 && (!q->ec_member_offsets || !q->ec_members ||
                                                                     !q->ec_sels || !q->ec_fks || !q->ec_vars))) {
        set_error("null array in query");
        return GQ_ERR_INVALID;
    }
    /* join widths are sums of base widths in an int (planner_info.cuh:22-28);
     * the device memo keeps them below 2^31 */
    long long wsum = 0;
    for (int i = 0; i < q->n_rels; i++) {
        if (q->base_rels[i].width < 0) {
            set_error("base relation %d: negative width", i);
            return GQ_ERR_INVALID;
        }
        wsum += q->base_rels[i].width;
    }
    if (wsum > 0x7fffffffLL) {
        set_error("base relation widths sum to %lld (at most 2^31-1)", wsum);
        return GQ_ERR_INVALID;
    }
    for (int e = 0; e < q->n_eq_classes; e++) {
        int b0 = q->ec_member_offsets[e], b1 = q->ec_member_offsets[e + 1];
        if (b1 - b0 < 1) {
            set_error("equivalence class %d has no members", e);
            return GQ_ERR_INVALID;
        }
        for (int m = b0; m < b1; m++) {
            if (q->ec_members[m] < 0 || q->ec_members[m] >= q->n_rels || (m > b0 && q->ec_members[m] <= q->ec_members[m - 1])) {
                set_error("equivalence class %d: members must be ascending relation indices", e);
                return GQ_ERR_INVALID;
            }
        }
    }
    return GQ_OK;
}

/* minimumSpanningTree (gpuqo_spanning_tree.cu:17-68): Prim's algorithm from
 * relation 0 on the join graph, each edge (i, j) weighted by the estimated
 * rows of joining base relations i and j over every class.  Each round takes
 * the first minimum in (i, j) order over i already in the tree and j outside
 * it, with the reference's strict `<` against FLT_MAX, so an edge whose rows
 * are +inf is never taken.  A NaN product of selectivity and rows is
 * clamped to 1 row by the row estimate (`rows > 1` is false for NaN), so
 * such an edge weighs 1 and is preferred.  Only the edge table changes: the classes,
 * and so every join cost, stay those of the full graph.  A graph in which no
 * edge can be taken keeps the tree built so far and is then rejected as
 * disconnected. */
template <int W>
static void spanning_tree(HQuery<W> &h) {
    HostView<W> hv;
    hv.build(h);
    std::vector<Set<W>> out(h.n);
    Set<W> in;
    in.set(0);
    for (int added = 1; added < h.n; added++) {
        float best = FLT_MAX;
        int bi = -1, bj = -1;
        for (int i = 0; i < h.n; i++) {
            if (!bit(in, i)) continue;
            for (int j = 0; j < h.n; j++) {
                if (!bit(h.adj[i], j) || bit(in, j)) continue;
                Set<W> l, r;
                l.set(i);
                r.set(j);
                uint64_t m[GQ_MAX_ECW];
                match_mask(hv.v, l, r, m);
                const float rows = join_rows(join_selectivity(hv.v, l, r, m), h.rows[i], h.rows[j]);
                if (rows < best) {
                    best = rows;
                    bi = i;
                    bj = j;
                }
            }
        }
        if (bi < 0) break;
        in.set(bj);
        out[bi].set(bj);
        out[bj].set(bi);
    }
    h.adj.swap(out);
}

template <int W>
static void convert(const gq_query *q, const gq_config *cfg, HQuery<W> &h) {  /* convertGpuqoPlannerInfo :133-209 */
    h.n = q->n_rels;
    h.n_iters = q->n_rels;
    h.p = make_params(cfg);
    h.rows.resize(h.n);
    h.tuples.resize(h.n);
    h.startup.resize(h.n);
    h.total.resize(h.n);
    h.width.resize(h.n);
    h.composite.assign(h.n, 0);
    h.adj.resize(h.n);
    for (int i = 0; i < h.n; i++) {
        h.rows[i] = q->base_rels[i].rows;
        h.tuples[i] = q->base_rels[i].tuples;
        h.startup[i] = q->base_rels[i].cost.startup;
        h.total[i] = q->base_rels[i].cost.total;
        h.width[i] = q->base_rels[i].width;
        h.adj[i] = set_from_words<W>(q->edge_table + (size_t)i * q->n_words, q->n_words);
        h.adj[i].unset(i);
    }
    int soff = 0;
    for (int e = 0; e < q->n_eq_classes; e++) {
        int b0 = q->ec_member_offsets[e], b1 = q->ec_member_offsets[e + 1];
        Set<W> ec;
        for (int m = b0; m < b1; m++) ec.set(q->ec_members[m]);
        h.ec_relids.push_back(ec);
        int s = b1 - b0;
        for (int j = 0; j < s * (s - 1) / 2; j++) h.sels.push_back(q->ec_sels[soff + j]);
        soff += s * (s - 1) / 2;
        for (int m = b0; m < b1; m++) {
            h.fks.push_back(set_from_words<W>(q->ec_fks + (size_t)m * q->n_words, q->n_words));
            h.vars.push_back(q->ec_vars[m]);
        }
    }
    if (cfg->gpuqo_spanning_tree) spanning_tree(h);
}

/* ------------------------------------------------------------ output */

template <int W>
static int emit(const QT<W> *t, gq_plan_out *out) {
    if (!t) return -1;
    if (out->n_nodes >= out->capacity) return -2;
    int me = out->n_nodes++;
    gq_plan_node &nd = out->nodes[me];
    memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < W && i < GQ_MAX_WORDS; i++) nd.relids[i] = t->id.w[i];
    nd.rows = t->rows;
    nd.cost.startup = t->startup;
    nd.cost.total = t->total;
    nd.width = t->width;
    int l = emit(t->left, out);
    if (l == -2) return -2;
    int r = emit(t->right, out);
    if (r == -2) return -2;
    out->nodes[me].left = l;
    out->nodes[me].right = r;
    return me;
}

template <int W>
static int plan_t(const gq_query *q, const gq_config *cfg, gq_plan_out *out, bool noalloc = false) {
    auto t0 = std::chrono::steady_clock::now();
    HQuery<W> h;
    convert(q, cfg, h);                /* host preparation first, then the device */
    if (int e = device_init(cfg->device)) return e;
    RunCtx ctx(cfg);
    ctx.noalloc = noalloc;
    QT<W> *t = nullptr;
    int rc = gpuqo_run_t(ctx, h, &t);
    out->n_nodes = 0;
    memset(&out->stats, 0, sizeof(out->stats));
    out->stats.candidates = ctx.stats.candidates;
    out->stats.sets = ctx.stats.sets;
    out->stats.pairs = ctx.stats.pairs;
    out->stats.memo_bytes = ctx.stats.memo_bytes;
    out->stats.gpu_ms = ctx.stats.gpu_ms;
    out->stats.wall_ms = host_now_ms() - std::chrono::duration<double, std::milli>(t0.time_since_epoch()).count();
    out->stats.n_dp_calls = ctx.dp_calls;
    out->stats.memo_kind = ctx.stats.memo_kind;
    out->stats.filter_ms = ctx.stats.filter_ms;
    out->stats.eval_ms = ctx.stats.eval_ms;
    out->stats.filter_launches = ctx.stats.filter_launches;
    out->stats.eval_launches = ctx.stats.eval_launches;
    if (rc != GQ_OK) return rc;
    int r = GQ_OK;
    if (!t) {
        set_error("no join tree (is the join graph connected?)");
        r = GQ_ERR_NO_PLAN;
    } else if (emit(t, out) == -2) {
        set_error("output capacity %d < %d nodes", out->capacity, 2 * q->n_rels - 1);
        r = GQ_ERR_CAPACITY;
    }
    const double t_end = host_now_ms();
    const double t_start = std::chrono::duration<double, std::milli>(t0.time_since_epoch()).count();
    if (ctx.stats.t_first_launch > 0.0) {
        out->stats.host_prep_ms = ctx.stats.t_first_launch - t_start;
        out->stats.host_enqueue_ms = ctx.stats.t_last_launch - ctx.stats.t_first_launch;
        out->stats.host_post_ms = t_end - ctx.stats.t_synced;
    }
    out->stats.wall_ms = t_end - t_start;
    return r;
}

/* ------------------------------------------------------------ snapshots */

/* A snapshot's breadth-first plan nodes as a pre-order gq_plan_out. */
static int emit_wide_nodes(const std::vector<WideNode> &nodes, gq_plan_out *out) {
    std::vector<QT<1>> made(nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) {
        made[i].id = Set<1>();
        made[i].id.w[0] = nodes[i].set;
        made[i].rows = nodes[i].rows;
        made[i].startup = nodes[i].startup;
        made[i].total = nodes[i].total;
        made[i].width = nodes[i].width;
        made[i].left = nodes[i].left >= 0 ? &made[nodes[i].left] : nullptr;
        made[i].right = nodes[i].right >= 0 ? &made[nodes[i].right] : nullptr;
    }
    out->n_nodes = 0;
    const int r = made.empty() ? -1 : emit(&made[0], out);
    if (r == -2) {
        set_error("output capacity %d below the plan's %zu nodes", out->capacity, nodes.size());
        return GQ_ERR_CAPACITY;
    }
    return r < 0 ? GQ_ERR_INTERNAL : GQ_OK;
}

static int plan_snapshot_impl(const gq_query *q, const gq_config *cfg, gq_plan_out *out, SnapTable *snap) {
    if (q->n_rels < 2 || q->n_rels > 64) {
        set_error("a memo snapshot takes one DP of 2..64 relations (got %d)", q->n_rels);
        return GQ_ERR_UNSUPPORTED;
    }
    if (cfg->gpuqo_algorithm != GQ_ALGO_DPSUB ||
        (cfg->gpuqo_idp_n_iters > 1 && cfg->gpuqo_idp_n_iters < q->n_rels)) {
        set_error("a memo snapshot takes one DPsub/MPDP DP (gpuqo_algorithm dpsub, no IDP)");
        return GQ_ERR_UNSUPPORTED;
    }
    /* the DP numbering: the BFS relabelling plan_t will apply */
    HQuery<1> h;
    convert(q, cfg, h);
    PlainPrepared pp;
    int rc = prepare_plain(h, pp);
    if (rc) return rc;
    SnapReq req;
    req.n = q->n_rels;
    memset(req.perm, -1, sizeof(req.perm));
    for (const auto &e : pp.bfs) req.perm[e.to] = (int8_t)ctz(e.from);
    g_snap_req = &req;
    rc = plan_t<1>(q, cfg, out);
    g_snap_req = nullptr;
    if (rc) {
        snapshot_free(req.out);
        return rc;
    }
    if (!req.done) {
        snapshot_free(req.out);
        set_error("no DP ran for the snapshot");
        return GQ_ERR_INTERNAL;
    }
    *snap = req.out;
    return GQ_OK;
}

}  // namespace gq

/* The opaque snapshot handle of the C ABI. */
struct gq_snapshot {
    gq::SnapTable t;
};

/* ================================================================ C ABI */

extern "C" {

gq_status gq_config_default(gq_config *c) {
    if (!c) return GQ_ERR_INVALID;
    memset(c, 0, sizeof(*c));
    c->gpuqo_algorithm = GQ_ALGO_DPSIZE;                 /* guc.c:4810-4816 */
    c->gpuqo_dpsub_bicc = 0;                             /* :1153-1161 */
    c->gpuqo_dpsub_filter = 1;                           /* :1113-1121 */
    c->gpuqo_dpsub_filter_threshold = 16;                /* :2190-2197 */
    c->gpuqo_dpsub_filter_cpu_enum_threshold = 1024;     /* :2200-2207 */
    c->gpuqo_dpsub_filter_keys_overprovisioning = 100;   /* :2210-2217 */
    c->gpuqo_dpsub_ccc = 1;                              /* :1123-1131 */
    c->gpuqo_dpsub_csg = 1;                              /* :1133-1141 */
    c->gpuqo_dpsub_csg_threshold = 128;                  /* :2220-2227 */
    c->gpuqo_dpsub_tree = 0;                             /* :1143-1151 */
    c->gpuqo_spanning_tree = 0;                          /* :1163-1171 */
    c->gpuqo_n_parallel = 10240;                         /* :2180-2187 */
    c->gpuqo_scratchpad_size_mb = 10;                    /* :2150-2157 */
    c->gpuqo_min_memo_size_mb = 1;                       /* :2170-2177 */
    c->gpuqo_max_memo_size_mb = 1024;                    /* :2160-2167 */
    c->gpuqo_idp_type = GQ_IDP2;                         /* :4819-4825 */
    c->gpuqo_idp_n_iters = 0;                            /* :2260-2267 */
    c->gpuqo_dpe_n_threads = 4;                          /* :2230-2237 */
    c->gpuqo_cpu_dpsub_parallel_chunk_size = 64;         /* :2240-2247 */
    c->gpuqo_dpe_pairs_per_depbuf = 128;                 /* :2250-2257 */
    c->seq_page_cost = 1.0;
    c->random_page_cost = 4.0;
    c->cpu_tuple_cost = 0.01;
    c->cpu_index_tuple_cost = 0.005;
    c->cpu_operator_cost = 0.0025;
    c->disable_cost = 1.0e10;
    c->effective_cache_size = 524288;
    c->work_mem = 4096;
    c->enable_seqscan = c->enable_indexscan = c->enable_tidscan = c->enable_sort = 1;
    c->enable_hashagg = c->enable_nestloop = c->enable_mergejoin = c->enable_hashjoin = 1;   /* guc.c:999-1027 */
    c->memo_kind = GQ_MEMO_AUTO;
    c->device = 0;
    return GQ_OK;
}

gq_status gq_init(int32_t device) {
    return gq::guarded([&] { return gq::device_init(device); });
}

static gq_status plan_impl(const gq_query *q, const gq_config *cfg, gq_plan_out *out) {
    int rc = gq::validate(q);
    if (rc) return (gq_status)rc;
    if (!cfg || !out || !out->nodes) {
        gq::set_error("null config or output");
        return GQ_ERR_INVALID;
    }
    if ((rc = gq::validate_config(cfg))) return (gq_status)rc;
    if (cfg->gpuqo_algorithm != GQ_ALGO_DPSUB && cfg->gpuqo_algorithm != GQ_ALGO_DPSIZE) {
        gq::set_error("algorithm %d is not a GPU algorithm (use dpsub or dpsize)", cfg->gpuqo_algorithm);
        return GQ_ERR_UNSUPPORTED;
    }
    if (q->n_rels <= 64) return (gq_status)gq::plan_t<1>(q, cfg, out);
    return (gq_status)gq::plan_t<4>(q, cfg, out);
}

gq_status gq_plan(const gq_query *q, const gq_config *cfg, gq_plan_out *out) {
    return gq::guarded([&] { return plan_impl(q, cfg, out); });
}

static gq_status plan_noalloc_impl(const gq_query *q, const gq_config *cfg, gq_plan_out *out) {
    int rc = gq::validate(q);
    if (rc) return (gq_status)rc;
    if (!cfg || !out || !out->nodes) {
        gq::set_error("null config or output");
        return GQ_ERR_INVALID;
    }
    if ((rc = gq::validate_config(cfg))) return (gq_status)rc;
    if (cfg->gpuqo_algorithm != GQ_ALGO_DPSUB) {
        gq::set_error("the allocation-free path plans GQ_ALGO_DPSUB (plain or bicc) only");
        return GQ_ERR_UNSUPPORTED;
    }
    if (q->n_rels > GQ_NOALLOC_MAX_RELS) {
        gq::set_error("the allocation-free path plans at most %d relations (got %d)", GQ_NOALLOC_MAX_RELS,
                      q->n_rels);
        return GQ_ERR_UNSUPPORTED;
    }
    if (cfg->gpuqo_idp_n_iters > 1 && cfg->gpuqo_idp_n_iters < q->n_rels) {
        gq::set_error("the allocation-free path runs no IDP decomposition");
        return GQ_ERR_UNSUPPORTED;
    }
    if (cfg->memo_kind != GQ_MEMO_AUTO && cfg->memo_kind != GQ_MEMO_DENSE) {
        gq::set_error("the allocation-free path uses the dense memo only");
        return GQ_ERR_UNSUPPORTED;
    }
    if (!gq::device_ready(cfg->device)) {
        gq::set_error("gq_init(%d) must be called before gq_plan_noalloc", cfg->device);
        return GQ_ERR_INVALID;
    }
    gq_config c = *cfg;
    c.profile_kernels = 0;
    c.gpuqo_max_memo_size_mb = std::max(c.gpuqo_max_memo_size_mb, 64);   /* 2^20 dense records fit */
    return (gq_status)gq::plan_t<1>(q, &c, out, true);
}

gq_status gq_plan_noalloc(const gq_query *q, const gq_config *cfg, gq_plan_out *out) {
    return gq::guarded([&] { return plan_noalloc_impl(q, cfg, out); });
}

static gq_status plan_batch_impl(int32_t n_queries, const gq_query *const *qs, const gq_config *cfg,
                                 gq_plan_out *outs) {
    if (n_queries < 0 || (n_queries > 0 && (!qs || !outs)) || !cfg) {
        gq::set_error("bad batch arguments");
        return GQ_ERR_INVALID;
    }
    int rc = gq::validate_config(cfg);
    if (rc) return (gq_status)rc;
    rc = gq::device_init(cfg->device);
    if (rc) return (gq_status)rc;
    /* Plain DPs (no IDP decomposition) that the fused kernel takes run in one
     * launch; every other query goes through gq_plan. */
    const bool plain_cfg = cfg->gpuqo_algorithm == GQ_ALGO_DPSUB &&
                           (cfg->gpuqo_idp_n_iters <= 1);
    std::vector<gq::PlainPrepared> preps(n_queries);
    std::vector<gq::DPInput> ins(n_queries);
    std::vector<const gq::DPInput *> batch;
    std::vector<gq::DPAlgo> algos;
    std::vector<int> where;
    const gq::DPAlgo algo = cfg->gpuqo_dpsub_bicc ? gq::DP_BICC : gq::DP_DPSUB;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n_queries; i++) {
        if ((rc = gq::validate(qs[i]))) return (gq_status)rc;
        if (!outs[i].nodes) {
            gq::set_error("null output nodes for query %d", i);
            return GQ_ERR_INVALID;
        }
    }
    /* host preparation of the plain DPs, in parallel over queries */
    std::vector<char> ok(n_queries, 0);
    gq::parallel_for(n_queries, [&](int i) {
        if (!plain_cfg || qs[i]->n_rels > 30 || qs[i]->n_rels < 2) return;
        gq::HQuery<1> h;
        gq::convert(qs[i], cfg, h);
        if (gq::prepare_plain(h, preps[i])) return;
        if (gq::build_dp_input(preps[i].ri, ins[i])) return;
        /* the lockstep batch keeps dense memos; others are planned alone */
        if (gq::resolve_memo(ins[i], cfg->memo_kind, cfg->gpuqo_max_memo_size_mb) != GQ_MEMO_DENSE) return;
        if (!gq::fused_eligible(ins[i], algo)) return;
        ok[i] = 1;
    });
    for (int i = 0; i < n_queries; i++) {
        if (!ok[i]) continue;
        batch.push_back(&ins[i]);
        algos.push_back(algo);
        where.push_back(i);
    }
    std::vector<char> done(n_queries, 0);
    if (!batch.empty()) {
        std::vector<std::vector<gq::DPNode>> trees;
        std::vector<gq::DPStats> sts;
        double ms = 0.0;
        if ((rc = gq::run_dp_batch(batch, algos, trees, sts, &ms))) return (gq_status)rc;
        auto t1 = std::chrono::steady_clock::now();
        double wall = std::chrono::duration<double, std::milli>(t1 - t0).count();
        std::vector<char> cap_err(batch.size(), 0);
        gq::parallel_for((int)batch.size(), [&](int b) {
            int i = where[b];
            gq::RunCtx ctx(cfg);
            gq::QT<1> *t = gq::finish_plain(ctx, preps[i], trees[b]);
            gq_plan_out *out = &outs[i];
            out->n_nodes = 0;
            memset(&out->stats, 0, sizeof(out->stats));
            if (gq::emit(t, out) == -2) cap_err[b] = 1;
        });
        for (size_t b = 0; b < batch.size(); b++) {
            int i = where[b];
            gq_plan_out *out = &outs[i];
            if (cap_err[b]) {
                gq::set_error("output capacity %d < %d nodes", out->capacity, 2 * qs[i]->n_rels - 1);
                return GQ_ERR_CAPACITY;
            }
            out->stats.candidates = sts[b].candidates;
            out->stats.sets = sts[b].sets;
            out->stats.pairs = sts[b].pairs;
            out->stats.memo_bytes = sts[b].memo_bytes;
            out->stats.gpu_ms = ms;            /* the whole batch launch */
            out->stats.wall_ms = wall;
            out->stats.n_dp_calls = 1;
            out->stats.memo_kind = sts[b].memo_kind;
            out->stats.eval_ms = ms;
            out->stats.eval_launches = 1;
            done[i] = 1;
        }
    }
    for (int i = 0; i < n_queries; i++) {
        if (done[i]) continue;
        gq_status st = plan_impl(qs[i], cfg, &outs[i]);
        if (st != GQ_OK) return st;
    }
    return GQ_OK;
}

gq_status gq_plan_batch(int32_t n_queries, const gq_query *const *qs, const gq_config *cfg, gq_plan_out *outs) {
    return gq::guarded([&] { return plan_batch_impl(n_queries, qs, cfg, outs); });
}

void gq_interrupt(int32_t on) { gq::g_interrupt.store(on ? 1 : 0); }

gq_status gq_tuning_get(gq_tuning *t) {
    if (!t) {
        gq::set_error("null tuning");
        return GQ_ERR_INVALID;
    }
    const gq::Tuning v = gq::tuning_get();
    t->tree_two_lane_waves = (int32_t)v.lpe2_waves;
    t->general_grid_max = (int32_t)v.gen_grid_max;
    t->tree_task_iters = (int32_t)v.tree_iters;
    t->tree_expand_max_rels = (int32_t)v.expand_max_n;
    t->tree_one_xcd_wgs = (int32_t)v.one_xcd_wgs;
    return GQ_OK;
}

static gq_status tuning_set_impl(const gq_tuning *t) {
    if (!t || t->tree_two_lane_waves < 1 || t->general_grid_max < 1 || t->tree_task_iters < 1 ||
        t->tree_expand_max_rels < 0 || t->tree_one_xcd_wgs < 0) {
        gq::set_error("null or non-positive tuning");
        return GQ_ERR_INVALID;
    }
    gq::Tuning v;
    v.lpe2_waves = (uint32_t)t->tree_two_lane_waves;
    v.gen_grid_max = (uint32_t)t->general_grid_max;
    v.tree_iters = (uint32_t)t->tree_task_iters;
    v.expand_max_n = (uint32_t)t->tree_expand_max_rels;
    v.one_xcd_wgs = (uint32_t)t->tree_one_xcd_wgs;
    return (gq_status)gq::tuning_set(v);
}

gq_status gq_tuning_set(const gq_tuning *t) {
    return gq::guarded([&] { return tuning_set_impl(t); });
}

static gq_status autotune_impl(int32_t n_queries, const gq_query *const *qs, const gq_config *cfg, int32_t reps,
                               int32_t knobs, gq_tuning *out) {
    if (n_queries < 1 || !qs || !cfg || reps < 1 || !out || (knobs & ~(GQ_TUNE_SINGLE | GQ_TUNE_BATCH))) {
        gq::set_error("bad autotune arguments");
        return GQ_ERR_INVALID;
    }
    int rc = gq::validate_config(cfg);
    if (rc) return (gq_status)rc;
    rc = gq::device_init(cfg->device);
    if (rc) return (gq_status)rc;
    std::vector<std::vector<gq_plan_node>> bufs(n_queries);
    std::vector<gq_plan_out> outs(n_queries);
    for (int i = 0; i < n_queries; i++) {
        if ((rc = gq::validate(qs[i]))) return (gq_status)rc;
        bufs[i].resize(2 * (size_t)qs[i]->n_rels);
        outs[i].nodes = bufs[i].data();
        outs[i].capacity = (int32_t)bufs[i].size();
    }
    /* batch knobs: wall time of gq_plan_batch (milliseconds, steady);
     * single-plan knobs: device time of each gq_plan summed over the sample
     * (a plan's wall time jitters by more than the knobs move it), with four
     * times the repetitions */
    auto median_ms = [&](bool batch, double *ms) -> int {
        std::vector<double> t;
        const int rr = batch ? reps : 4 * reps;
        for (int r = 0; r <= rr; r++) {                      /* the first run warms up */
            const auto t0 = std::chrono::steady_clock::now();
            double dev = 0.0;
            if (batch) {
                int e = plan_batch_impl(n_queries, qs, cfg, outs.data());
                if (e) return e;
            } else {
                for (int i = 0; i < n_queries; i++) {
                    int e = plan_impl(qs[i], cfg, &outs[i]);
                    if (e) return e;
                    dev += outs[i].stats.gpu_ms;
                }
            }
            const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (r) t.push_back(batch ? wall : dev);
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return GQ_OK;
    };
    gq::Tuning best = gq::tuning_get();
    struct Knob { uint32_t gq::Tuning::*field; bool batch; std::vector<uint32_t> cands; };
    const Knob knob_list[] = {
        {&gq::Tuning::lpe2_waves, false, {1024, 2048, 4096, 8192, 16384}},
        {&gq::Tuning::expand_max_n, false, {0, 16, 18, 20, 22}},
        {&gq::Tuning::one_xcd_wgs, false, {0, 8, 32, 64}},
        {&gq::Tuning::tree_iters, true, {4, 8, 16, 32}},
        {&gq::Tuning::gen_grid_max, true, {256, 512, 1024, 2048}},
    };
    /* a knob leaves its current value only for a candidate at least 1%
     * faster, so timing noise does not move it */
    for (const Knob &kb : knob_list) {
        if (!(knobs & (kb.batch ? GQ_TUNE_BATCH : GQ_TUNE_SINGLE))) continue;
        const uint32_t cur = best.*kb.field;
        double cur_ms = 0.0;
        if ((rc = median_ms(kb.batch, &cur_ms))) return (gq_status)rc;
        uint32_t pick = cur;
        double pick_ms = cur_ms;
        for (uint32_t v : kb.cands) {
            if (v == cur) continue;
            gq::Tuning t = best;
            t.*kb.field = v;
            if ((rc = gq::tuning_set(t))) return (gq_status)rc;
            double ms = 0.0;
            if ((rc = median_ms(kb.batch, &ms))) {
                gq::tuning_set(best);
                return (gq_status)rc;
            }
            if (ms < pick_ms && ms < 0.99 * cur_ms) {
                pick = v;
                pick_ms = ms;
            }
        }
        best.*kb.field = pick;
        if ((rc = gq::tuning_set(best))) return (gq_status)rc;
    }
    return gq_tuning_get(out);
}

gq_status gq_autotune(int32_t n_queries, const gq_query *const *qs, const gq_config *cfg, int32_t reps,
                      int32_t knobs, gq_tuning *out) {
    const gq::Tuning before = gq::tuning_get();
    gq_status st = gq::guarded([&] { return autotune_impl(n_queries, qs, cfg, reps, knobs, out); });
    if (st != GQ_OK) gq::tuning_set(before);     /* a failed search leaves the knobs as they were */
    return st;
}

const char *gq_last_error(void) { return gq::g_err; }

const char *gq_version(void) { return "gpuqo-mi355x 0.1 (gfx950)"; }

/* ------------------------------------------------------------ snapshots */

gq_status gq_plan_snapshot(const gq_query *q, const gq_config *cfg, gq_plan_out *out, gq_snapshot **snap) {
    return gq::guarded([&]() -> int {
        if (!snap) {
            gq::set_error("null snapshot pointer");
            return GQ_ERR_INVALID;
        }
        *snap = nullptr;
        int rc = gq::validate(q);
        if (rc) return rc;
        if (!cfg || !out || !out->nodes) {
            gq::set_error("null config or output");
            return GQ_ERR_INVALID;
        }
        if ((rc = gq::validate_config(cfg))) return rc;
        gq::SnapTable t;
        rc = gq::plan_snapshot_impl(q, cfg, out, &t);
        if (rc) return rc;
        gq_snapshot *h = new (std::nothrow) gq_snapshot;
        if (!h) {
            gq::snapshot_free(t);
            gq::set_error("host allocation failed");
            return GQ_ERR_NOMEM;
        }
        h->t = t;
        *snap = h;
        return GQ_OK;
    });
}

gq_status gq_snapshot_subplan(const gq_snapshot *snap, const uint64_t *relids, gq_plan_out *out) {
    return gq::guarded([&]() -> int {
        if (!snap || !relids || !out || !out->nodes) {
            gq::set_error("null snapshot, relation set or output");
            return GQ_ERR_INVALID;
        }
        std::vector<gq::WideNode> nodes;
        int rc = gq::snapshot_walk(snap->t, relids[0], nodes);
        if (rc) return rc;
        memset(&out->stats, 0, sizeof(out->stats));
        return gq::emit_wide_nodes(nodes, out);
    });
}

gq_status gq_snapshot_info(const gq_snapshot *snap, int32_t *n_rels, uint64_t *n_sets, uint64_t *device_bytes) {
    if (!snap) {
        gq::set_error("null snapshot");
        return GQ_ERR_INVALID;
    }
    if (n_rels) *n_rels = snap->t.n_rels;
    if (n_sets) *n_sets = snap->t.n_sets;
    if (device_bytes) *device_bytes = snap->t.tab ? (32ull << snap->t.lg) : 0ull;
    return GQ_OK;
}

gq_status gq_snapshot_export(const gq_snapshot *snap, void *buf, uint64_t cap, uint64_t *bytes) {
    return gq::guarded([&]() -> int {
        if (!snap || !bytes) {
            gq::set_error("null snapshot or size pointer");
            return GQ_ERR_INVALID;
        }
        const uint64_t need = GQ_SNAPSHOT_HEADER_BYTES + GQ_SNAPSHOT_RECORD_BYTES * snap->t.n_sets;
        *bytes = need;
        if (!buf) return GQ_OK;                 /* size query */
        if (cap < need) {
            gq::set_error("export buffer of %llu bytes, %llu needed", (unsigned long long)cap,
                          (unsigned long long)need);
            return GQ_ERR_CAPACITY;
        }
        std::vector<gq::SnapRecord> recs;
        int rc = gq::snapshot_read(snap->t, recs);
        if (rc) return rc;
        if (recs.size() != snap->t.n_sets) {
            gq::set_error("snapshot holds %zu records, %llu expected", recs.size(),
                          (unsigned long long)snap->t.n_sets);
            return GQ_ERR_INTERNAL;
        }
        uint8_t *p = (uint8_t *)buf;
        memset(p, 0, GQ_SNAPSHOT_HEADER_BYTES);
        memcpy(p, GQ_SNAPSHOT_MAGIC, 8);
        const uint32_t version = GQ_SNAPSHOT_VERSION, nr = (uint32_t)snap->t.n_rels,
                       rb = GQ_SNAPSHOT_RECORD_BYTES;
        const uint64_t ns = recs.size();
        memcpy(p + 8, &version, 4);
        memcpy(p + 12, &nr, 4);
        memcpy(p + 16, &ns, 8);
        memcpy(p + 24, &rb, 4);
        uint8_t *r = p + GQ_SNAPSHOT_HEADER_BYTES;
        for (const gq::SnapRecord &x : recs) {
            memcpy(r, &x.set, 8);
            memcpy(r + 8, &x.left, 8);
            memcpy(r + 16, &x.rows, 4);
            memcpy(r + 20, &x.startup, 4);
            memcpy(r + 24, &x.total, 4);
            memcpy(r + 28, &x.width, 4);
            r += GQ_SNAPSHOT_RECORD_BYTES;
        }
        return GQ_OK;
    });
}

gq_status gq_snapshot_import(const void *buf, uint64_t bytes, gq_snapshot **snap) {
    return gq::guarded([&]() -> int {
        if (!buf || !snap) {
            gq::set_error("null buffer or snapshot pointer");
            return GQ_ERR_INVALID;
        }
        *snap = nullptr;
        const uint8_t *p = (const uint8_t *)buf;
        uint32_t version = 0, nr = 0, rb = 0;
        uint64_t ns = 0;
        if (bytes < GQ_SNAPSHOT_HEADER_BYTES || memcmp(p, GQ_SNAPSHOT_MAGIC, 8) != 0) {
            gq::set_error("not a memo snapshot (magic)");
            return GQ_ERR_INVALID;
        }
        memcpy(&version, p + 8, 4);
        memcpy(&nr, p + 12, 4);
        memcpy(&ns, p + 16, 8);
        memcpy(&rb, p + 24, 4);
        if (version != GQ_SNAPSHOT_VERSION || rb != GQ_SNAPSHOT_RECORD_BYTES || nr < 1 || nr > 64 || ns < nr ||
            ns > (1ull << 30) || bytes != GQ_SNAPSHOT_HEADER_BYTES + GQ_SNAPSHOT_RECORD_BYTES * ns) {
            gq::set_error("memo snapshot header: version %u, %u relations, %llu records of %u bytes in %llu bytes",
                          version, nr, (unsigned long long)ns, rb, (unsigned long long)bytes);
            return GQ_ERR_INVALID;
        }
        const uint64_t all = nr >= 64 ? ~0ull : ((1ull << nr) - 1ull);
        std::vector<gq::SnapRecord> recs((size_t)ns);
        const uint8_t *r = p + GQ_SNAPSHOT_HEADER_BYTES;
        for (uint64_t i = 0; i < ns; i++, r += GQ_SNAPSHOT_RECORD_BYTES) {
            gq::SnapRecord &x = recs[i];
            memcpy(&x.set, r, 8);
            memcpy(&x.left, r + 8, 8);
            memcpy(&x.rows, r + 16, 4);
            memcpy(&x.startup, r + 20, 4);
            memcpy(&x.total, r + 24, 4);
            memcpy(&x.width, r + 28, 4);
            const bool base = __builtin_popcountll(x.set) == 1;
            const bool left_ok = base ? x.left == 0ull : (x.left && !(x.left & ~x.set) && x.left != x.set);
            if (!x.set || (x.set & ~all) || x.width < 0 || !left_ok) {
                gq::set_error("memo snapshot record %llu is malformed", (unsigned long long)i);
                return GQ_ERR_INVALID;
            }
        }
        gq::SnapTable t;
        int rc = gq::snapshot_build(recs, (int)nr, t);
        if (rc) return rc;
        gq_snapshot *h = new (std::nothrow) gq_snapshot;
        if (!h) {
            gq::snapshot_free(t);
            gq::set_error("host allocation failed");
            return GQ_ERR_NOMEM;
        }
        h->t = t;
        *snap = h;
        return GQ_OK;
    });
}

gq_status gq_snapshot_free(gq_snapshot *snap) {
    return gq::guarded([&]() -> int {
        if (!snap) return GQ_OK;
        gq::snapshot_free(snap->t);
        delete snap;
        return GQ_OK;
    });
}

}  // extern "C"
