/*
 * gq_server.cpp -- the continuous-batching planning server (gq_server_*).
 *
 * No counterpart in the reference, which plans one query per backend call
 * (gpuqo_main_internal.cu:177-185).  Three kinds of threads:
 *
 *   - the caller submits requests (a deep copy of the query is kept) and
 *     polls finished ones;
 *   - `max_inflight` request threads each run gq_plan on an admitted
 *     request -- plain DP, IDP2 or UnionDP, unchanged -- with their GPU DPs
 *     routed (DPExecutor) to the engine instead of launched directly;
 *   - one engine thread owns the StepEngine: it admits pending DPs into its
 *     device arena, advances every active DP by one level per step, and hands
 *     finished plans back to the waiting request threads.
 *
 * Requests are admitted in arrival order while the projected memo bytes of
 * the running ones plus the next fit the budget (a request is always
 * admitted when nothing runs).  Requests that arrive while others are on the
 * device join them at the next engine step; a UnionDP request's partition
 * DPs share steps with everyone else's DPs, and the engine's device arena
 * persists across its iterations.  A request's plan equals gq_plan's for the
 * same query: the engine runs the same kernels on the same inputs.
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gpuqo_amd.h"
#include "gq_dp.h"
#include "gq_internal.h"

namespace gq {

namespace {

using Clock = std::chrono::steady_clock;

/* A deep copy of a gq_query. */
struct OwnedQuery {
    std::vector<gq_base_rel> base;
    std::vector<uint64_t> edges, fks;
    std::vector<int32_t> moff, mem;
    std::vector<float> sels;
    std::vector<gq_var_info> vars;
    gq_query q;

    void copy(const gq_query *s) {
        const int n = s->n_rels, nw = s->n_words, ne = s->n_eq_classes;
        base.assign(s->base_rels, s->base_rels + n);
        edges.assign(s->edge_table, s->edge_table + (size_t)n * nw);
        moff.assign(s->ec_member_offsets, s->ec_member_offsets + ne + 1);
        const int nm = moff[ne];
        mem.assign(s->ec_members, s->ec_members + nm);
        size_t ns = 0;
        for (int e = 0; e < ne; e++) {
            const size_t sz = (size_t)(moff[e + 1] - moff[e]);
            ns += sz * (sz - 1) / 2;
        }
        sels.assign(s->ec_sels, s->ec_sels + ns);
        fks.assign(s->ec_fks, s->ec_fks + (size_t)nm * nw);
        vars.assign(s->ec_vars, s->ec_vars + nm);
        q = *s;
        q.base_rels = base.data();
        q.edge_table = edges.data();
        q.ec_member_offsets = moff.data();
        q.ec_members = mem.data();
        q.ec_sels = sels.data();
        q.ec_fks = fks.data();
        q.ec_vars = vars.data();
    }
};

struct Request {
    Request *qnext = nullptr;          /* link in the one queue holding it */
    uint64_t tag = 0;
    gq_config cfg;
    OwnedQuery oq;
    uint64_t projected = 0;
    Clock::time_point t_submit;
    std::vector<gq_plan_node> nodes;
    gq_plan_stats stats;
    gq_status st = GQ_OK;
};

/* FIFO of requests linked through Request::qnext: moving a request between
 * the waiting, admitted and finished queues allocates nothing, so a request
 * thread can always hand back its result. */
struct ReqQueue {
    Request *head = nullptr, *tail = nullptr;
    size_t n = 0;
    bool empty() const { return n == 0; }
    size_t size() const { return n; }
    Request *front() const { return head; }
    void push_back(Request *r) {
        r->qnext = nullptr;
        if (tail) tail->qnext = r;
        else head = r;
        tail = r;
        n++;
    }
    void pop_front() {
        head = head->qnext;
        if (!head) tail = nullptr;
        n--;
    }
    void delete_all() {
        while (head) {
            Request *r = head;
            pop_front();
            delete r;
        }
    }
};

/* One run_many call of a request thread: its DPs are in the engine's hands
 * until `remaining` drops to zero.  `refs` counts the queue and engine
 * entries that still point at it; the thread waits on its own `cv` until
 * the work is done and unreferenced, so a step wakes only the threads whose
 * DPs it finished. */
struct DPWork {
    const std::vector<const DPInput *> *ins;
    const std::vector<DPAlgo> *algos;
    std::vector<std::vector<DPNode>> *trees;
    std::vector<DPStats> *sts;
    size_t remaining = 0;
    size_t refs = 0;
    int rc = GQ_OK;
    bool done = false;
    std::condition_variable cv;
};

}  // namespace

struct Server;

struct ServerExecutor : DPExecutor {
    Server *srv = nullptr;
    int run(const std::vector<const DPInput *> &ins, const std::vector<DPAlgo> &algos,
            std::vector<std::vector<DPNode>> &trees, std::vector<DPStats> &sts) override;
};

struct Server {
    gq_config cfg;
    uint64_t budget = 0;
    int max_inflight = 1;
    StepEngine *eng = nullptr;

    std::mutex mu;
    std::condition_variable cv_engine, cv_work, cv_done;
    ReqQueue waiting, admitted, finished;
    int running = 0;
    uint64_t running_bytes = 0;
    size_t outstanding = 0;
    bool stop = false;

    std::deque<std::pair<DPWork *, size_t>> dp_pending;
    std::unordered_map<uint64_t, std::pair<DPWork *, size_t>> in_engine;
    uint64_t next_key = 1;

    std::thread engine_th;
    std::vector<std::thread> workers;

    /* with mu held */
    void admit_waiting() {
        while (!waiting.empty()) {
            Request *r = waiting.front();
            const bool fits = running == 0 || (running < max_inflight && running_bytes + r->projected <= budget);
            if (!fits) break;
            waiting.pop_front();
            running++;
            running_bytes += r->projected;
            admitted.push_back(r);
            cv_work.notify_one();
        }
    }

    /* with mu held: fail every DP of `w` that is not finished */
    static void fail_work(DPWork *w, int rc) {
        w->rc = rc;
        w->remaining = 0;
        w->done = true;
    }
    /* with mu held: one queue or engine entry of `w` is gone */
    static void unref(DPWork *w) {
        if (--w->refs == 0 && w->done) w->cv.notify_one();
    }

    void engine_loop() {
        (void)hipSetDevice(cfg.device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_engine.wait(lk, [&] { return stop || !dp_pending.empty() || engine_active(eng) > 0; });
            if (dp_pending.empty() && engine_active(eng) == 0) {
                if (stop) break;
                continue;
            }
            while (!dp_pending.empty()) {
                DPWork *w = dp_pending.front().first;
                const size_t i = dp_pending.front().second;
                if (w->done) {                       /* failed earlier: drop */
                    dp_pending.pop_front();
                    unref(w);
                    continue;
                }
                const uint64_t key = next_key++;
                bool admitted_dp = false;
                const gq_status ast = guarded([&] {
                    in_engine.reserve(in_engine.size() + 1);
                    admitted_dp = engine_admit(eng, *(*w->ins)[i], (*w->algos)[i], key);
                    if (admitted_dp) in_engine[key] = std::make_pair(w, i);
                    return (int)GQ_OK;
                });
                if (ast != GQ_OK) {                  /* host allocation failed: fail this DP's work */
                    fail_work(w, ast);
                    dp_pending.pop_front();
                    unref(w);
                    continue;
                }
                if (admitted_dp) {
                    dp_pending.pop_front();
                    continue;
                }
                if (engine_active(eng) == 0) {       /* can never fit the arena */
                    set_error("a DP of %d relations does not fit the memo budget", (*w->ins)[i]->hdr.n_rels);
                    fail_work(w, GQ_ERR_NOMEM);
                    dp_pending.pop_front();
                    unref(w);
                    continue;
                }
                break;
            }
            if (engine_active(eng) == 0) continue;
            lk.unlock();
            std::vector<EngineDone> done;
            const int rc = guarded([&] { return engine_step(eng, done); });
            lk.lock();
            for (EngineDone &d : done) {
                auto it = in_engine.find(d.key);
                if (it == in_engine.end()) continue;
                DPWork *w = it->second.first;
                const size_t i = it->second.second;
                in_engine.erase(it);
                if (!w->done) {
                    if (d.rc) {
                        fail_work(w, d.rc);
                    } else {
                        (*w->trees)[i] = std::move(d.tree);
                        (*w->sts)[i] = d.st;
                        if (--w->remaining == 0) w->done = true;
                    }
                }
                unref(w);
            }
            if (rc) {                                /* device failure: fail everything on the device */
                std::vector<DPWork *> ws;
                for (auto &kv : in_engine) ws.push_back(kv.second.first);
                for (auto &p : dp_pending) ws.push_back(p.first);
                in_engine.clear();
                dp_pending.clear();
                for (DPWork *w : ws) fail_work(w, rc);
                for (DPWork *w : ws) unref(w);
                engine_destroy(eng);
                eng = nullptr;
                if (guarded([&] { return engine_create(cfg.device, budget, &eng); }) != GQ_OK) break;
            }
        }
    }

    void worker_loop() {
        (void)hipSetDevice(cfg.device);
        ServerExecutor ex;
        ex.srv = this;
        thread_executor() = &ex;
        for (;;) {
            Request *r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !admitted.empty(); });
                if (admitted.empty()) break;        /* stop */
                r = admitted.front();
                admitted.pop_front();
            }
            gq_plan_out out;
            out.nodes = r->nodes.data();
            out.capacity = (int32_t)r->nodes.size();
            out.n_nodes = 0;
            memset(&out.stats, 0, sizeof(out.stats));
            r->st = gq_plan(&r->oq.q, &r->cfg, &out);
            r->nodes.resize(out.n_nodes > 0 ? out.n_nodes : 0);   /* shrinks: no allocation */
            r->stats = out.stats;
            r->stats.wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
            {
                std::lock_guard<std::mutex> g(mu);
                running--;
                running_bytes -= r->projected;
                finished.push_back(r);
                admit_waiting();
                cv_done.notify_all();
            }
        }
        thread_executor() = nullptr;
    }
};

int ServerExecutor::run(const std::vector<const DPInput *> &ins, const std::vector<DPAlgo> &algos,
                        std::vector<std::vector<DPNode>> &trees, std::vector<DPStats> &sts) {
    trees.assign(ins.size(), std::vector<DPNode>());
    sts.assign(ins.size(), DPStats());
    if (ins.empty()) return GQ_OK;
    DPWork w;
    w.ins = &ins;
    w.algos = &algos;
    w.trees = &trees;
    w.sts = &sts;
    w.remaining = ins.size();
    w.refs = 0;
    std::unique_lock<std::mutex> lk(srv->mu);
    if (!srv->eng) {
        set_error("planning server engine unavailable");
        return GQ_ERR_DEVICE;
    }
    /* each queued entry is a reference; if queueing fails part way, the
     * entries already queued are dropped by the engine thread (w is done) */
    const gq_status qst = guarded([&] {
        for (size_t i = 0; i < ins.size(); i++) {
            srv->dp_pending.push_back(std::make_pair(&w, i));
            w.refs++;
        }
        return (int)GQ_OK;
    });
    if (qst != GQ_OK) Server::fail_work(&w, qst);
    srv->cv_engine.notify_one();
    /* DPs of a failed work may still be queued or on the device: wait until
     * the engine no longer references this stack object too */
    w.cv.wait(lk, [&] { return w.done && w.refs == 0; });
    return w.rc;
}

/* Connected sets of a tree-shaped query graph (0 when the graph is not a
 * tree): sum over vertices v of P_v(1), P_v(x) = x prod_c (1 + P_c(x)). */
static uint64_t tree_set_count(const gq_query *q) {
    const int n = q->n_rels, nw = q->n_words;
    if (n < 1 || n > 32) return 0;
    std::vector<uint32_t> adj(n, 0u);
    uint64_t deg = 0;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++)
            if (j != i && ((q->edge_table[(size_t)i * nw + j / 64] >> (j % 64)) & 1ull)) adj[i] |= 1u << j;
        deg += (uint64_t)__builtin_popcount(adj[i]);
    }
    if (deg != 2ull * (uint64_t)(n - 1)) return 0;
    std::vector<int> order(1, 0), parent(n, -1);
    uint32_t seen = 1u;
    for (size_t x = 0; x < order.size(); x++) {
        const int v = order[x];
        uint32_t nb = adj[v] & ~seen;
        while (nb) {
            const int u = __builtin_ctz(nb);
            nb &= nb - 1u;
            seen |= 1u << u;
            parent[u] = v;
            order.push_back(u);
        }
    }
    if ((int)order.size() != n) return 0;
    std::vector<uint64_t> f(n, 1ull);          /* f(v) = P_v(1) */
    uint64_t total = 0;
    for (int x = n - 1; x >= 0; x--) {
        const int v = order[x];
        total += f[v];
        if (parent[v] >= 0) f[parent[v]] *= 1ull + f[v];
    }
    return total;
}

/* Memo bytes a request will hold at most at one time on the engine. */
static uint64_t projected_bytes(const gq_query *q, const gq_config *c) {
    const int n = q->n_rels;
    const int k = c->gpuqo_idp_n_iters;
    const double max_memo = (double)c->gpuqo_max_memo_size_mb * 1048576.0;
    if (k <= 1 || k >= n) {
        const int nd = std::min(n, GQ_MAX_DP_RELS - 2);
        const uint64_t dense = engine_slice_bytes(nd);
        const uint64_t sets = tree_set_count(q);
        const bool tree = sets && c->gpuqo_algorithm == GQ_ALGO_DPSUB;
        if (tree && c->memo_kind == GQ_MEMO_HASH) return engine_hash_slice_bytes(sets);
        if (tree && (c->memo_kind == GQ_MEMO_RANK ||
                     (c->memo_kind == GQ_MEMO_AUTO && 36.0 * (double)(1ull << nd) > max_memo)))
            return engine_rank_slice_bytes(sets);
        return dense;
    }
    /* IDP2 / UnionDP: DPs of at most k relations (dense up to the memo
     * limit, rank or open-addressing above it), plus a level of small UnionDP
     * partitions (fewer than 16 relations each) batched together */
    const uint64_t dp = std::min<uint64_t>(engine_slice_bytes(std::min(k, GQ_MAX_DP_RELS - 2)),
                                           (uint64_t)max_memo + (64ull << 20));
    return dp + 32 * engine_slice_bytes(15);
}

}  // namespace gq

struct gq_server {
    gq::Server s;
};

namespace gq {

/* Stop a server's threads and free it (the threads in `started`). */
static void server_shutdown(gq_server *srv) {
    Server &S = srv->s;
    {
        std::lock_guard<std::mutex> g(S.mu);
        S.stop = true;
        S.outstanding -= S.waiting.size();
        S.waiting.delete_all();
    }
    S.cv_work.notify_all();
    for (auto &t : S.workers) t.join();
    {
        std::lock_guard<std::mutex> g(S.mu);
        S.cv_engine.notify_all();
    }
    if (S.engine_th.joinable()) S.engine_th.join();
    S.finished.delete_all();
    S.admitted.delete_all();
    engine_destroy(S.eng);
    delete srv;
}

static gq_status server_create_impl(const gq_config *cfg, uint64_t memo_budget_bytes, int32_t max_inflight,
                                    gq_server **out) {
    if (!out || !cfg || max_inflight < 1 || memo_budget_bytes == 0) {
        set_error("bad server arguments");
        return GQ_ERR_INVALID;
    }
    if (int rc = validate_config(cfg)) return (gq_status)rc;
    *out = nullptr;
    std::unique_ptr<gq_server> owner(new gq_server());
    Server &S = owner->s;
    S.cfg = *cfg;
    S.budget = memo_budget_bytes;
    S.max_inflight = max_inflight;
    int rc = engine_create(cfg->device, memo_budget_bytes, &S.eng);
    if (rc) {
        engine_destroy(S.eng);
        return (gq_status)rc;
    }
    gq_server *srv = owner.release();
    try {
        S.workers.reserve((size_t)max_inflight);
        S.engine_th = std::thread([&S] { S.engine_loop(); });
        for (int i = 0; i < max_inflight; i++) S.workers.emplace_back([&S] { S.worker_loop(); });
    } catch (...) {
        server_shutdown(srv);            /* joins the threads that did start */
        throw;
    }
    *out = srv;
    return GQ_OK;
}

static gq_status server_submit_impl(gq_server *srv, const gq_query *q, const gq_config *cfg, uint64_t tag) {
    if (!srv || !q) {
        set_error("bad submit arguments");
        return GQ_ERR_INVALID;
    }
    Server &S = srv->s;
    const gq_config *c = cfg ? cfg : &S.cfg;
    if (c->device != S.cfg.device) {
        set_error("request device %d differs from the server's %d", c->device, S.cfg.device);
        return GQ_ERR_INVALID;
    }
    if (int rc = validate(q)) return (gq_status)rc;
    if (int rc = validate_config(c)) return (gq_status)rc;
    std::unique_ptr<Request> r(new Request());
    r->tag = tag;
    r->cfg = *c;
    r->oq.copy(q);
    r->projected = projected_bytes(q, c);
    r->nodes.resize((size_t)2 * q->n_rels);
    memset(&r->stats, 0, sizeof(r->stats));
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(S.mu);
    if (S.stop) {
        set_error("server is stopping");
        return GQ_ERR_INVALID;
    }
    if (S.waiting.size() >= 65536) {
        set_error("server queue full");
        return GQ_ERR_BUSY;
    }
    S.waiting.push_back(r.release());
    S.outstanding++;
    S.admit_waiting();
    return GQ_OK;
}

static gq_status server_poll_impl(gq_server *srv, int32_t wait, int32_t *done, uint64_t *tag,
                                  gq_status *req_status, gq_plan_out *out) {
    if (!srv || !done || !out) {
        set_error("bad poll arguments");
        return GQ_ERR_INVALID;
    }
    Server &S = srv->s;
    *done = 0;
    std::unique_lock<std::mutex> lk(S.mu);
    auto ready = [&] { return !S.finished.empty() || S.outstanding == 0; };
    if (wait < 0) S.cv_done.wait(lk, ready);
    else if (wait > 0) S.cv_done.wait_for(lk, std::chrono::milliseconds(wait), ready);
    if (S.finished.empty()) return GQ_OK;
    Request *r = S.finished.front();
    if (out->capacity < (int32_t)r->nodes.size() || (!r->nodes.empty() && !out->nodes)) {
        set_error("output capacity %d < %zu nodes", out->capacity, r->nodes.size());
        return GQ_ERR_CAPACITY;
    }
    S.finished.pop_front();
    S.outstanding--;
    lk.unlock();
    if (!r->nodes.empty()) memcpy(out->nodes, r->nodes.data(), sizeof(gq_plan_node) * r->nodes.size());
    out->n_nodes = (int32_t)r->nodes.size();
    out->stats = r->stats;
    if (tag) *tag = r->tag;
    if (req_status) *req_status = r->st;
    *done = 1;
    delete r;
    return GQ_OK;
}

}  // namespace gq

extern "C" {

gq_status gq_server_create(const gq_config *cfg, uint64_t memo_budget_bytes, int32_t max_inflight,
                           gq_server **out) {
    return gq::guarded([&] { return gq::server_create_impl(cfg, memo_budget_bytes, max_inflight, out); });
}

gq_status gq_server_submit_cfg(gq_server *srv, const gq_query *q, const gq_config *cfg, uint64_t tag) {
    return gq::guarded([&] { return gq::server_submit_impl(srv, q, cfg, tag); });
}

gq_status gq_server_submit(gq_server *srv, const gq_query *q, uint64_t tag) {
    return gq_server_submit_cfg(srv, q, nullptr, tag);
}

gq_status gq_server_poll(gq_server *srv, int32_t wait, int32_t *done, uint64_t *tag, gq_status *req_status,
                         gq_plan_out *out) {
    return gq::guarded([&] { return gq::server_poll_impl(srv, wait, done, tag, req_status, out); });
}

gq_status gq_server_destroy(gq_server *srv) {
    if (!srv) return GQ_OK;
    return gq::guarded([&] {
        gq::server_shutdown(srv);
        return (int)GQ_OK;
    });
}

}  // extern "C"
