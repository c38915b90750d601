/*
 * gq_drivers.h -- BFS relabelling, the Remapper, and the DP / IDP1 / IDP2 /
 * UnionDP drivers (host side).
 */
#ifndef GQ_DRIVERS_H
#define GQ_DRIVERS_H

#include <deque>
#include <memory>
#include <vector>

#include "gq_host.h"

namespace gq {

/* Per-call state: configuration, counters, and an arena owning every plan
 * node built during the call (trees share composite subtrees). */
struct RunCtx {
    int algorithm, bicc, idp_type, idp_n_iters, memo_kind, profile, max_memo_mb;
    bool noalloc = false;      /* gq_plan_noalloc: DPs in the static scratch array */
    DPStats stats;
    int dp_calls = 0;
    std::deque<QT<1>> arena1;
    std::deque<QT<4>> arena4;
    explicit RunCtx(const gq_config *cfg)
        : algorithm(cfg->gpuqo_algorithm), bicc(cfg->gpuqo_dpsub_bicc), idp_type(cfg->gpuqo_idp_type),
          idp_n_iters(cfg->gpuqo_idp_n_iters), memo_kind(cfg->memo_kind), profile(cfg->profile_kernels),
          max_memo_mb(cfg->gpuqo_max_memo_size_mb) {}
    template <int W>
    QT<W> *node();
    void add(const DPStats &s) {
        stats.candidates += s.candidates;
        stats.sets += s.sets;
        stats.pairs += s.pairs;
        stats.gpu_ms += s.gpu_ms;
        stats.filter_ms += s.filter_ms;
        stats.eval_ms += s.eval_ms;
        stats.filter_launches += s.filter_launches;
        stats.eval_launches += s.eval_launches;
        if (s.memo_bytes > stats.memo_bytes) stats.memo_bytes = s.memo_bytes;
        stats.memo_kind = s.memo_kind;
        if (s.t_first_launch > 0.0 && (stats.t_first_launch == 0.0 || s.t_first_launch < stats.t_first_launch))
            stats.t_first_launch = s.t_first_launch;
        if (s.t_last_launch > stats.t_last_launch) stats.t_last_launch = s.t_last_launch;
        if (s.t_synced > stats.t_synced) stats.t_synced = s.t_synced;
    }
};
template <>
inline QT<1> *RunCtx::node<1>() { arena1.emplace_back(); return &arena1.back(); }
template <>
inline QT<4> *RunCtx::node<4>() { arena4.emplace_back(); return &arena4.back(); }

/* __gpuqo_run (gpuqo_main_internal.cu:81-147). *out is arena-owned. */
template <int W>
int gpuqo_run_t(RunCtx &ctx, const HQuery<W> &h, QT<W> **out);

int build_dp_input(const HQuery<1> &q, DPInput &in);

/* Plain-DP pieces for batching: relabel a query in BFS order, and map a DP
 * tree (in the relabelled numbering) back to the caller's. */
struct PlainPrepared {
    HQuery<1> ri;
    std::vector<RemapEl<1>> bfs;
};
int prepare_plain(const HQuery<1> &h, PlainPrepared &p);
QT<1> *finish_plain(RunCtx &ctx, const PlainPrepared &p, const std::vector<DPNode> &nodes);

}  // namespace gq

#endif
