/*
 * gq_drivers.cpp -- relabelling and the large-query drivers.
 *
 *   BFS relabel      gpuqo_bfs_indexing.cu:13-55
 *   Remapper         gpuqo_remapper.cu:17-344
 *   dispatch         gpuqo_main_internal.cu:19-147
 *   IDP2 + GOO       gpuqo_idp.cu:95-297, gpuqo_cpu_goo.cu:39-148
 *   UnionDP          gpuqo_dpdp_union.cu:35-472
 *
 * Relation i is bit i here (the reference uses bit i+1); every order the
 * drivers depend on (numeric set order, lowest member first, BFS with
 * ascending neighbours, libstdc++ priority_queue / push_heap on the same
 * comparators) is unchanged by the shift.
 */
#include <algorithm>
#include <cstring>
#include <map>
#include <queue>
#include <vector>

#include "gq_drivers.h"
#include "gq_wide.h"

namespace gq {

/* ------------------------------------------------------------ Remapper */

template <int WI, int WO>
struct Remapper {
    std::vector<RemapEl<WI>> tr;
    /* every entry maps one relation and none carries a subtree (a pure
     * relabelling, such as the BFS one): the maps below then take a table
     * lookup per member instead of a pass over the entries, with the same
     * results */
    bool perm = true;
    int16_t to_of[64 * WI], from_of[64 * WO];
    explicit Remapper(const std::vector<RemapEl<WI>> &t) : tr(t) {
        for (auto &x : to_of) x = -1;
        for (auto &x : from_of) x = -1;
        for (const auto &e : tr) {
            if (popc(e.from) != 1 || e.qt) {
                perm = false;
                break;
            }
            to_of[ctz(e.from)] = (int16_t)e.to;
            from_of[e.to] = (int16_t)ctz(e.from);
        }
    }

    Set<WO> fwd(const Set<WI> &id) const {
        Set<WO> o;
        if (perm) {
            for (Set<WI> t = id; !empty(t);) {
                const int f = ctz(t);
                t.unset(f);
                if (to_of[f] >= 0) o.set(to_of[f]);
            }
            return o;
        }
        for (const auto &e : tr)
            if (intersects(e.from, id)) o.set(e.to);
        return o;
    }
    Set<WO> fwd_plain(const Set<WI> &id, const HQuery<WI> &old) const {   /* remapRelidNoComposite */
        Set<WO> o;
        if (perm) {
            for (Set<WI> t = id; !empty(t);) {
                const int f = ctz(t);
                t.unset(f);
                if (to_of[f] >= 0 && !old.composite[f]) o.set(to_of[f]);
            }
            return o;
        }
        for (const auto &e : tr) {
            if (popc(e.from) > 1 || old.composite[ctz(e.from)]) continue;
            if (intersects(e.from, id)) o.set(e.to);
        }
        return o;
    }
    Set<WI> inv(const Set<WO> &id) const {
        Set<WI> o;
        if (perm) {
            for (Set<WO> t = id; !empty(t);) {
                const int v = ctz(t);
                t.unset(v);
                if (from_of[v] >= 0) o.set(from_of[v]);
            }
            return o;
        }
        for (const auto &e : tr)
            if (bit(id, e.to)) o |= e.from;
        return o;
    }
    bool dropped(const Set<WI> &ec) const {
        if (perm) {                                 /* ec within one mapped relation */
            const int c = popc(ec);
            return c == 0 ? !tr.empty() : (c == 1 && to_of[ctz(ec)] >= 0);
        }
        for (const auto &e : tr)
            if (subset_of(ec, e.from)) return true;
        return false;
    }
    static Set<WO> nth_member(int idx, const Set<WO> &s) {
        Set<WO> t = s;
        for (int i = 0; i < idx; i++) t.unset(ctz(t));
        return lowbit(t);
    }

    /* remapPlannerInfo (gpuqo_remapper.cu:200-276) */
    void remap(const HQuery<WI> &old, HQuery<WO> &h) const {
        HostView<WI> ov;                            /* built when a merged class needs it */
        bool ov_built = false;
        const int n = (int)tr.size();
        h = HQuery<WO>();
        h.n = n;
        h.n_iters = old.n_iters;
        h.p = old.p;
        h.rows.assign(n, 0.0f);
        h.tuples.assign(n, 0.0f);
        h.startup.assign(n, 0.0f);
        h.total.assign(n, 0.0f);
        h.width.assign(n, 0);
        h.composite.assign(n, 0);
        h.adj.assign(n, Set<WO>());
        for (const auto &e : tr) {                      /* remapEdgeTable :84-107 */
            Set<WO> a;
            Set<WI> t = e.from;
            while (!empty(t)) {
                int f = ctz(t);
                t.unset(f);
                a |= fwd(old.adj[f]);
            }
            a.unset(e.to);
            h.adj[e.to] = a;
        }
        for (const auto &e : tr) {                      /* remapBaseRels :109-140 */
            int o = ctz(e.from);
            if (e.qt) {
                h.rows[e.to] = e.qt->rows;
                h.startup[e.to] = e.qt->startup;
                h.total[e.to] = e.qt->total;
                h.width[e.to] = e.qt->width;
                h.tuples[e.to] = e.qt->rows;
            } else {
                h.rows[e.to] = old.rows[o];
                h.startup[e.to] = old.startup[o];
                h.total[e.to] = old.total[o];
                h.width[e.to] = old.width[o];
                h.tuples[e.to] = old.tuples[o];
            }
            h.composite[e.to] = popc(e.from) == 1 ? old.composite[o] : 1;
        }
        int old_moff = 0, old_soff = 0;                 /* remapEqClass :142-198 */
        for (size_t i = 0; i < old.ec_relids.size(); i++) {
            const Set<WI> &ecf = old.ec_relids[i];
            int s_from = popc(ecf);
            if (!dropped(ecf)) {
                Set<WO> ect = fwd(ecf);
                int s_to = popc(ect);
                size_t sb = h.sels.size(), mb = h.fks.size();
                h.sels.resize(sb + s_to * (s_to - 1) / 2, 0.0f);
                h.fks.resize(mb + s_to, Set<WO>());
                h.vars.resize(mb + s_to);
                for (int lt = 0; lt < s_to; lt++) {
                    Set<WO> lto = nth_member(lt, ect);
                    Set<WI> lfr = inv(lto);
                    int lf = rank_below(ecf, lfr);
                    for (int rt = lt + 1; rt < s_to; rt++) {
                        Set<WO> rto = nth_member(rt, ect);
                        Set<WI> rfr = inv(rto);
                        int rf = rank_below(ecf, rfr);
                        float sel;
                        if (popc(lfr) == popc(lto) && popc(rfr) == popc(rto)) {
                            sel = old.sels[old_soff + eq_class_index(lf, rf, s_from)];
                        } else {
                            if (!ov_built) {
                                ov.build(old);
                                ov_built = true;
                            }
                            sel = ec_selectivity(ov.v, (int)i, lfr, rfr);
                        }
                        h.sels[sb + eq_class_index(lt, rt, s_to)] = sel;
                    }
                    if (popc(lfr) == 1 && !old.composite[ctz(lfr)])
                        h.fks[mb + lt] = fwd_plain(old.fks[old_moff + lf], old);
                    h.vars[mb + lt] = old.vars[old_moff + lf];
                    if (popc(lfr) > 1) h.vars[mb + lt].index_available = 0;
                }
                h.ec_relids.push_back(ect);
            }
            old_moff += s_from;
            old_soff += s_from * (s_from - 1) / 2;
        }
    }

    /* remapQueryTree (:278-301) */
    QT<WI> *back(RunCtx &ctx, const QT<WO> *t) const {
        if (!t) return nullptr;
        if (popc(t->id) == 1 && !perm) {
            int idx = ctz(t->id);
            for (const auto &e : tr)
                if (e.qt && e.to == idx) return e.qt;
        }
        QT<WI> *o = ctx.node<WI>();
        o->id = inv(t->id);
        o->left = back(ctx, t->left);
        o->right = back(ctx, t->right);
        o->rows = t->rows;
        o->startup = t->startup;
        o->total = t->total;
        o->width = t->width;
        return o;
    }

    /* remapQueryTreeFwd (:303-321) */
    QT<WO> *forward(RunCtx &ctx, const QT<WI> *t) const {
        if (!t) return nullptr;
        QT<WO> *o = ctx.node<WO>();
        o->id = fwd(t->id);
        if (popc(o->id) == 1) {
            o->left = o->right = nullptr;
        } else {
            o->left = forward(ctx, t->left);
            o->right = forward(ctx, t->right);
        }
        o->rows = t->rows;
        o->startup = t->startup;
        o->total = t->total;
        o->width = t->width;
        return o;
    }
};

/* makeBFSIndexRemapper (gpuqo_bfs_indexing.cu:13-55) */
template <int W>
static std::vector<RemapEl<W>> bfs_list(const HQuery<W> &h) {
    std::vector<int> q;
    q.reserve(h.n);
    q.push_back(0);
    size_t front = 0;
    int idx = 0;
    std::vector<RemapEl<W>> out;
    Set<W> seen = Set<W>::nth(0);
    while (front < q.size() && idx < h.n) {
        int u = q[front++];
        out.push_back(RemapEl<W>{Set<W>::nth(u), idx++, nullptr});
        Set<W> edges = h.adj[u];
        while (!empty(edges)) {
            int v = ctz(edges);
            edges.unset(v);
            if (!bit(seen, v)) q.push_back(v);
        }
        seen |= h.adj[u];
    }
    return out;
}

template <int W>
static std::vector<RemapEl<W>> identity_list(const HQuery<W> &h) {
    std::vector<RemapEl<W>> out;
    for (int i = 0; i < h.n; i++) out.push_back(RemapEl<W>{Set<W>::nth(i), i, nullptr});
    return out;
}

template <int W>
static std::vector<RemapEl<W>> members_list(const Set<W> &s) {   /* relations of s in ascending order */
    std::vector<RemapEl<W>> out;
    Set<W> t = s;
    int j = 0;
    while (!empty(t)) {
        int i = ctz(t);
        t.unset(i);
        out.push_back(RemapEl<W>{Set<W>::nth(i), j++, nullptr});
    }
    return out;
}

/* ------------------------------------------------------------ the DP */

/* gpuqo_run_switch (gpuqo_main_internal.cu:19-59): the GPU algorithms. */
template <int W>
static int run_switch(RunCtx &ctx, const HQuery<W> &h, QT<W> **out);

template <>
int run_switch<4>(RunCtx &, const HQuery<4> &h, QT<4> **out) {
    *out = nullptr;
    set_error("a single DP over %d relations is not supported (max %d)", h.n, GQ_MAX_DP_RELS);
    return GQ_ERR_UNSUPPORTED;
}

static QT<1> *nodes_to_tree(RunCtx &ctx, const std::vector<DPNode> &nodes) {
    std::vector<QT<1> *> made(nodes.size(), nullptr);
    for (size_t i = 0; i < nodes.size(); i++) {
        QT<1> *t = ctx.node<1>();
        t->id.w[0] = nodes[i].set;
        t->rows = nodes[i].rows;
        t->startup = nodes[i].startup;
        t->total = nodes[i].total;
        t->width = nodes[i].width;
        made[i] = t;
    }
    for (size_t i = 0; i < nodes.size(); i++) {
        made[i]->left = nodes[i].left >= 0 ? made[nodes[i].left] : nullptr;
        made[i]->right = nodes[i].right >= 0 ? made[nodes[i].right] : nullptr;
    }
    return made.empty() ? nullptr : made[0];
}

static QT<1> *wide_nodes_to_tree(RunCtx &ctx, const std::vector<WideNode> &nodes) {
    std::vector<QT<1> *> made(nodes.size(), nullptr);
    for (size_t i = 0; i < nodes.size(); i++) {
        QT<1> *t = ctx.node<1>();
        t->id.w[0] = nodes[i].set;
        t->rows = nodes[i].rows;
        t->startup = nodes[i].startup;
        t->total = nodes[i].total;
        t->width = nodes[i].width;
        made[i] = t;
    }
    for (size_t i = 0; i < nodes.size(); i++) {
        made[i]->left = nodes[i].left >= 0 ? made[nodes[i].left] : nullptr;
        made[i]->right = nodes[i].right >= 0 ? made[nodes[i].right] : nullptr;
    }
    return made.empty() ? nullptr : made[0];
}

/* One DP over 33..64 relations (a tree join graph): the wide DP. */
static int run_wide(RunCtx &ctx, const HQuery<1> &h, DPAlgo algo, QT<1> **out) {
    *out = nullptr;
    if (algo == DP_DPSIZE) {
        set_error("DPsize runs on the GPU for at most %d relations (got %d)", GQ_MAX_DP_RELS, h.n);
        return GQ_ERR_UNSUPPORTED;
    }
    if (h.n_iters != h.n) {
        set_error("IDP1's level-capped DPs run on the GPU for at most %d relations (got %d)", GQ_MAX_DP_RELS, h.n);
        return GQ_ERR_UNSUPPORTED;
    }
    if (ctx.noalloc) {
        set_error("the allocation-free path plans at most %d relations", GQ_NOALLOC_MAX_RELS);
        return GQ_ERR_UNSUPPORTED;
    }
    WideInput wi;
    int rc = build_wide_input(h, wi);
    if (rc) return rc;
    std::vector<WideNode> nodes;
    DPStats st;
    rc = run_wide_dp(wi, algo, ctx.max_memo_mb, nodes, st);
    ctx.dp_calls++;
    if (rc) return rc;
    ctx.add(st);
    *out = wide_nodes_to_tree(ctx, nodes);
    return GQ_OK;
}

static QT<1> *single_leaf(RunCtx &ctx, const HQuery<1> &h) {
    QT<1> *t = ctx.node<1>();
    t->id = Set<1>::nth(0);
    t->rows = h.rows[0];
    t->startup = h.startup[0];
    t->total = h.total[0];
    base_cost(h.p, h.tuples[0], h.composite[0] != 0, &t->startup, &t->total);   /* cost_baserel */
    t->width = h.width[0];
    return t;
}

static DPAlgo dp_algo(const RunCtx &ctx) {
    return ctx.algorithm == GQ_ALGO_DPSIZE ? DP_DPSIZE : (ctx.bicc ? DP_BICC : DP_DPSUB);
}

/* Run the DPs of several relabelled queries: those the fused kernel takes
 * go through one batched launch, the others one by one. */
static int run_many(RunCtx &ctx, const std::vector<const HQuery<1> *> &hs, std::vector<QT<1> *> &out) {
    out.assign(hs.size(), nullptr);
    if (interrupted()) {
        set_error("interrupted");
        return GQ_ERR_INTERRUPTED;
    }
    const DPAlgo algo = dp_algo(ctx);
    DPExecutor *ex = thread_executor();
    std::vector<DPInput> ins(hs.size());
    std::vector<const DPInput *> batch;
    std::vector<DPAlgo> algos;
    std::vector<size_t> where;
    for (size_t i = 0; i < hs.size(); i++) {
        if (hs[i]->n == 1) {
            out[i] = single_leaf(ctx, *hs[i]);
            continue;
        }
        if (hs[i]->n > GQ_MAX_DP_RELS) {
            int rc = run_wide(ctx, *hs[i], algo, &out[i]);
            if (rc) return rc;
            continue;
        }
        int rc = build_dp_input(*hs[i], ins[i]);
        if (rc) return rc;
        if (algo != DP_DPSIZE) prepare_memo(ins[i], ctx.memo_kind, ctx.max_memo_mb);
        const bool to_batch = ex ? true : (hs.size() > 1 && ins[i].memo == GQ_MEMO_DENSE);
        if (to_batch && hs[i]->n_iters == hs[i]->n && fused_eligible(ins[i], algo)) {
            batch.push_back(&ins[i]);
            algos.push_back(algo);
            where.push_back(i);
            continue;
        }
        std::vector<DPNode> nodes;
        DPStats st;
        rc = run_dp(ins[i], algo, hs[i]->n_iters, ctx.memo_kind, ctx.profile != 0, nodes, st, ctx.noalloc);
        ctx.dp_calls++;
        if (rc) return rc;
        ctx.add(st);
        out[i] = nodes_to_tree(ctx, nodes);
    }
    if (!batch.empty()) {
        std::vector<std::vector<DPNode>> trees;
        std::vector<DPStats> sts;
        double ms = 0.0;
        int rc = ex ? ex->run(batch, algos, trees, sts) : run_dp_batch(batch, algos, trees, sts, &ms);
        if (rc) return rc;
        for (size_t b = 0; b < batch.size(); b++) {
            ctx.dp_calls++;
            ctx.add(sts[b]);
            out[where[b]] = nodes_to_tree(ctx, trees[b]);
        }
    }
    return GQ_OK;
}

template <>
int run_switch<1>(RunCtx &ctx, const HQuery<1> &h, QT<1> **out) {
    std::vector<QT<1> *> o;
    int rc = run_many(ctx, {&h}, o);
    *out = rc ? nullptr : o[0];
    return rc;
}

/* Relabel `lst` out of `info` and run the DP on the result
 * (gpuqo_run_idp2_dp idp.cu:122-141, gpuqo_run_dpdp_union_dp union.cu:261-280). */
template <int WO, int WI>
static int sub_dp(RunCtx &ctx, const HQuery<WO> &info, const std::vector<RemapEl<WO>> &lst, QT<WO> **out) {
    Remapper<WO, WI> rm(lst);
    HQuery<WI> ni;
    rm.remap(info, ni);
    ni.n_iters = ni.n;
    QT<WI> *t = nullptr;
    int rc = run_switch<WI>(ctx, ni, &t);
    if (rc) return rc;
    *out = rm.back(ctx, t);
    return GQ_OK;
}

/* ------------------------------------------------------------ GOO */

template <int W>
static bool valid_pair(const HQuery<W> &h, const Set<W> &l, const Set<W> &r) {   /* are_valid_pair, filter.cuh:60-90 */
    Set<W> nb;
    Set<W> t = l;
    while (!empty(t)) {
        int i = ctz(t);
        t.unset(i);
        nb |= h.adj[i];
    }
    return intersects(nb, r) && !intersects(l, r);
}

template <int W>
static float tree_join_rows(const HostView<W> &hv, const QT<W> &L, const QT<W> &R) {
    uint64_t m[GQ_MAX_ECW];
    match_mask(hv.v, L.id, R.id, m);
    return join_rows(join_selectivity(hv.v, L.id, R.id, m), L.rows, R.rows);
}

/* gpuqo_cpu_goo (gpuqo_cpu_goo.cu:59-148): greedy on result rows, GOO cost
 * = rows + children's costs, startup 0. */
template <int W>
static QT<W> *goo(RunCtx &ctx, const HQuery<W> &h) {
    HostView<W> hv;
    hv.build(h);
    auto cmp = [](const QT<W> *l, const QT<W> *r) {
        if (l && r) return l->rows > r->rows;
        return r != nullptr;
    };
    auto join = [&](QT<W> &L, QT<W> &R) {
        QT<W> *t = ctx.node<W>();
        t->id = L.id | R.id;
        t->rows = tree_join_rows(hv, L, R);
        t->startup = 0.0f;
        t->total = t->rows + L.total + R.total;
        t->left = &L;
        t->right = &R;
        t->width = L.width + R.width;
        return t;
    };
    std::vector<QT<W> *> heap;
    std::map<Set<W>, QT<W> *> rels;
    for (int i = 0; i < h.n; i++) {
        QT<W> *t = ctx.node<W>();
        t->id = Set<W>::nth(i);
        t->rows = h.rows[i];
        t->startup = 0.0f;
        t->total = h.rows[i];
        t->width = h.width[i];
        rels.insert(std::make_pair(t->id, t));
    }
    for (auto i = rels.begin(); i != rels.end(); ++i)
        for (auto j = i; j != rels.end(); ++j)
            if (valid_pair(h, i->first, j->first)) {
                heap.push_back(join(*i->second, *j->second));
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
    QT<W> *smallest = nullptr;
    if (h.n == 1) return rels.begin()->second;
    do {
        if (heap.empty()) return nullptr;
        std::pop_heap(heap.begin(), heap.end(), cmp);
        smallest = heap.back();
        heap.pop_back();
        auto lp = rels.find(smallest->left->id), rp = rels.find(smallest->right->id);
        if (lp == rels.end() || rp == rels.end()) {
            smallest = nullptr;
            continue;
        }
        rels.erase(lp);
        rels.erase(rp);
        for (auto j = rels.begin(); j != rels.end(); ++j)
            if (valid_pair(h, smallest->id, j->first)) {
                heap.push_back(join(*smallest, *j->second));
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
        rels.insert(std::make_pair(smallest->id, smallest));
    } while (smallest == nullptr || popc(smallest->id) < h.n);
    return smallest;
}

/* ------------------------------------------------------------ IDP2 */

template <int W>
static QT<W> *most_expensive_subtree(QT<W> *t, int max_size) {     /* idp.cu:95-120 */
    if (popc(t->id) <= max_size) return t;
    QT<W> *l = most_expensive_subtree(t->left, max_size);
    QT<W> *r = most_expensive_subtree(t->right, max_size);
    if (popc(l->id) == 1) return r;
    if (popc(r->id) == 1) return l;
    if (l->total > r->total) return l;
    if (l->total < r->total) return r;
    if (popc(l->id) < popc(r->id)) return l;
    return r;
}

template <int WO, int WI>
static int idp2_rec(RunCtx &ctx, QT<WO> *goo_t, const HQuery<WO> &info, const std::vector<RemapEl<WO>> &lst,
                    int n_iters, QT<WO> **out) {                                   /* idp.cu:143-257 */
    Remapper<WO, WI> rm(lst);
    HQuery<WI> ni;
    rm.remap(info, ni);
    QT<WI> *ngoo = rm.forward(ctx, goo_t);
    ni.n_iters = std::min(ni.n, n_iters);
    QT<WI> *maxt = most_expensive_subtree(ngoo, ni.n_iters);
    std::vector<RemapEl<WI>> rl = members_list(maxt->id);
    QT<WI> *reopt = nullptr;
    int rc;
    if (rl.size() < 64) rc = sub_dp<WI, 1>(ctx, ni, rl, &reopt);
    else rc = sub_dp<WI, 4>(ctx, ni, rl, &reopt);
    if (rc) return rc;
    QT<WI> *res = nullptr;
    if (ni.n_iters == ni.n) {
        res = reopt;
    } else {
        std::vector<RemapEl<WI>> nl;
        nl.push_back(RemapEl<WI>{reopt->id, 0, reopt});
        int j = 1;
        for (int k = 0; k < ni.n; k++)
            if (!subset_of(Set<WI>::nth(k), reopt->id)) nl.push_back(RemapEl<WI>{Set<WI>::nth(k), j++, nullptr});
        if (WI == 1 || nl.size() < 64) rc = idp2_rec<WI, 1>(ctx, ngoo, ni, nl, n_iters, &res);
        else rc = idp2_rec<WI, 4>(ctx, ngoo, ni, nl, n_iters, &res);
        if (rc) return rc;
    }
    *out = rm.back(ctx, res);
    return GQ_OK;
}

template <int W>
static int run_idp2(RunCtx &ctx, const HQuery<W> &h, QT<W> **out) {              /* idp.cu:259-297 */
    QT<W> *g = goo(ctx, h);
    if (!g) {
        set_error("GOO found no join tree (disconnected graph?)");
        return GQ_ERR_NO_PLAN;
    }
    return idp2_rec<W, W>(ctx, g, h, identity_list(h), ctx.idp_n_iters, out);
}

/* ------------------------------------------------------------ IDP1 */

template <int W>
static int run_idp1(RunCtx &ctx, HQuery<W> &info, QT<W> **out);

/* gpuqo_run_idp1_next (gpuqo_idp.cu:15-36) */
template <int WO, int WI>
static int idp1_next(RunCtx &ctx, const HQuery<WO> &info, const std::vector<RemapEl<WO>> &lst, QT<WO> **out) {
    Remapper<WO, WI> rm(lst);
    HQuery<WI> ni;
    rm.remap(info, ni);
    QT<WI> *t = nullptr;
    int rc = run_idp1<WI>(ctx, ni, &t);
    if (rc) return rc;
    *out = rm.back(ctx, t);
    return GQ_OK;
}

/* IDP1 (gpuqo_idp.cu:38-80): a DP capped at n_iters levels returns the
 * cheapest plan of n_iters relations (run_dp), which becomes relation 0 of
 * the next query, the remaining relations following in index order, until
 * one DP covers the whole query. */
template <int W>
static int run_idp1(RunCtx &ctx, HQuery<W> &info, QT<W> **out) {
    info.n_iters = std::min(info.n, ctx.idp_n_iters);
    QT<W> *qt = nullptr;
    int rc = run_switch<W>(ctx, info, &qt);
    if (rc) return rc;
    if (info.n_iters == info.n) {
        *out = qt;
        return GQ_OK;
    }
    std::vector<RemapEl<W>> lst;
    lst.push_back(RemapEl<W>{qt->id, 0, qt});
    int j = 1;
    for (int i = 0; i < info.n; i++)
        if (!subset_of(Set<W>::nth(i), qt->id)) lst.push_back(RemapEl<W>{Set<W>::nth(i), j++, nullptr});
    if (W == 1 || lst.size() < 64) return idp1_next<W, 1>(ctx, info, lst, out);
    return idp1_next<W, 4>(ctx, info, lst, out);
}

/* ------------------------------------------------------------ UnionDP */

struct GraphEdge {                   /* union.cu:111-123 */
    int left, right;                 /* relation indices */
    int left_size, right_size, total_size;
    float weight;                    /* rows = selectivity * rows_l * rows_r */
};

struct CompareEdges {                /* union.cu:147-155 */
    bool operator()(const GraphEdge *l, const GraphEdge *r) const {
        if (l->total_size == r->total_size) return l->weight > r->weight;
        return l->total_size > r->total_size;
    }
};

template <int WO, int WI>
static int union_rec(RunCtx &ctx, const HQuery<WO> &info, const std::vector<RemapEl<WO>> &lst, int n_iters,
                     QT<WO> **out) {                                               /* union.cu:283-444 */
    Remapper<WO, WI> rm(lst);
    HQuery<WI> ni;
    rm.remap(info, ni);
    ni.n_iters = std::min(ni.n, n_iters);
    if (ni.n == ni.n_iters) return sub_dp<WO, WI>(ctx, info, lst, out);
    HostView<WI> hv;
    hv.build(ni);
    /* fillPriorityQueues (:213-259): BFS from relation 0; an edge is made for
     * every neighbour not yet expanded, weighted by its join rows. */
    std::vector<GraphEdge> edges;
    edges.reserve(4 * ni.n);
    {
        std::queue<int> bq;
        int bfs_idx = 0;
        bq.push(0);
        Set<WI> seen2 = Set<WI>::nth(0);
        while (!bq.empty() && bfs_idx < ni.n) {
            int u = bq.front();
            bq.pop();
            Set<WI> es = ni.adj[u];
            bfs_idx++;
            while (!empty(es)) {
                int v = ctz(es);
                es.unset(v);
                if (!bit(seen2, v)) {
                    bq.push(v);
                    Set<WI> l = Set<WI>::nth(u), r = Set<WI>::nth(v);
                    uint64_t m[GQ_MAX_ECW];
                    match_mask(hv.v, l, r, m);
                    float sel = join_selectivity(hv.v, l, r, m);
                    GraphEdge e;
                    e.left = u;
                    e.right = v;
                    e.left_size = e.right_size = 1;
                    e.total_size = 2;
                    e.weight = sel * ni.rows[u] * ni.rows[v];
                    edges.push_back(e);
                }
            }
            seen2.set(u);
        }
    }
    std::priority_queue<GraphEdge *, std::vector<GraphEdge *>, CompareEdges> pq;
    for (auto &e : edges) pq.push(&e);
    /* DisjointSet (:35-109) over relation indices */
    std::vector<int> parent(ni.n), size(ni.n, 1);
    std::vector<Set<WI>> csg(ni.n);
    for (int i = 0; i < ni.n; i++) {
        parent[i] = i;
        csg[i] = Set<WI>::nth(i);
    }
    auto find = [&](int x) {
        int r = x;
        while (parent[r] != r) r = parent[r];
        while (parent[x] != r) {
            int nx = parent[x];
            parent[x] = r;
            x = nx;
        }
        return r;
    };
    const int upper_threshold = 16;                                  /* :320 */
    while (!pq.empty()) {                                            /* :324-344 */
        GraphEdge *e = pq.top();
        pq.pop();
        int x = find(e->left), y = find(e->right);
        if (x == y) continue;
        if (e->total_size != size[x] + size[y]) {
            e->left_size = size[x];
            e->right_size = size[y];
            e->total_size = e->left_size + e->right_size;
            pq.push(e);
        } else if (e->total_size < upper_threshold) {
            if (size[x] > size[y]) { parent[y] = x; size[x] += size[y]; csg[x] |= csg[y]; }
            else if (size[x] < size[y]) { parent[x] = y; size[y] += size[x]; csg[y] |= csg[x]; }
            else { parent[x] = y; size[y] *= 2; csg[y] |= csg[x]; }
        }
    }
    std::vector<Set<WI>> parts;                                      /* :357-367 */
    Set<WI> seen;
    for (int i = 0; i < ni.n; i++) {
        Set<WI> g = csg[find(i)];
        if (!subset_of(g, seen)) parts.push_back(g);
        seen |= g;
    }
    /* the partitions' DPs are independent: relabel each (:386-419) and run
     * them together */
    std::vector<RemapEl<WI>> next;
    std::vector<Remapper<WI, 1>> rms;
    std::vector<HQuery<1>> subs(parts.size());
    rms.reserve(parts.size());
    for (size_t i = 0; i < parts.size(); i++) {
        rms.emplace_back(members_list(parts[i]));
        rms.back().remap(ni, subs[i]);
        subs[i].n_iters = subs[i].n;
    }
    std::vector<const HQuery<1> *> hs;
    for (auto &x : subs) hs.push_back(&x);
    std::vector<QT<1> *> res1;
    int rc0 = run_many(ctx, hs, res1);
    if (rc0) return rc0;
    for (size_t i = 0; i < parts.size(); i++) {
        QT<WI> *r = rms[i].back(ctx, res1[i]);
        next.push_back(RemapEl<WI>{r->id, (int)i, r});
    }
    QT<WI> *res = nullptr;
    int rc;
    if (WI == 1 || next.size() < 64) rc = union_rec<WI, 1>(ctx, ni, next, n_iters, &res);
    else rc = union_rec<WI, 4>(ctx, ni, next, n_iters, &res);
    if (rc) return rc;
    *out = rm.back(ctx, res);
    return GQ_OK;
}

template <int W>
static int run_union(RunCtx &ctx, const HQuery<W> &h, QT<W> **out) {             /* union.cu:447-472 */
    return union_rec<W, W>(ctx, h, identity_list(h), ctx.idp_n_iters, out);
}

/* ------------------------------------------------------------ entry */

template <int W>
int gpuqo_run_t(RunCtx &ctx, const HQuery<W> &h, QT<W> **out) {
    *out = nullptr;
    std::vector<RemapEl<W>> bfs = bfs_list(h);
    if ((int)bfs.size() != h.n) {
        set_error("join graph is disconnected (%zu of %d relations reachable)", bfs.size(), h.n);
        return GQ_ERR_NO_PLAN;
    }
    Remapper<W, W> rm(bfs);
    HQuery<W> ri;
    rm.remap(h, ri);
    QT<W> *t = nullptr;
    int rc;
    if (ctx.idp_n_iters <= 1 || ctx.idp_n_iters >= h.n) {
        rc = run_switch<W>(ctx, ri, &t);
    } else if (ctx.idp_type == GQ_IDP2) {
        rc = run_idp2(ctx, ri, &t);
    } else if (ctx.idp_type == GQ_UNION) {
        rc = run_union(ctx, ri, &t);
    } else {
        rc = run_idp1(ctx, ri, &t);
    }
    if (rc) return rc;
    *out = rm.back(ctx, t);
    return GQ_OK;
}

template int gpuqo_run_t<1>(RunCtx &, const HQuery<1> &, QT<1> **);

int prepare_plain(const HQuery<1> &h, PlainPrepared &p) {
    p.bfs = bfs_list(h);
    if ((int)p.bfs.size() != h.n) {
        set_error("join graph is disconnected (%zu of %d relations reachable)", p.bfs.size(), h.n);
        return GQ_ERR_NO_PLAN;
    }
    Remapper<1, 1> rm(p.bfs);
    rm.remap(h, p.ri);
    return GQ_OK;
}

QT<1> *finish_plain(RunCtx &ctx, const PlainPrepared &p, const std::vector<DPNode> &nodes) {
    Remapper<1, 1> rm(p.bfs);
    return rm.back(ctx, nodes_to_tree(ctx, nodes));
}
template int gpuqo_run_t<4>(RunCtx &, const HQuery<4> &, QT<4> **);

}  // namespace gq
