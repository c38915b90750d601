/*
 * gq_dp.h -- host-side interface of the GPU DP (DPsub / MPDP / DPsize) over
 * one query of at most 32 relations, already relabelled.  Internal to the
 * library; the public C ABI is include/gpuqo_amd.h.
 */
#ifndef GQ_DP_H
#define GQ_DP_H

#include <stdint.h>
#include <vector>

#include "gq_cost_model.h"

namespace gq {

/* Flattened query for the device (32-bit relation sets, relation i = bit i).
 * Layout of the blob (32-bit words) described by BlobHdr; staged into LDS by
 * every kernel. */
struct BlobHdr {
    int n_rels, n_ec, n_mem, n_sels, ecw;
    Params p;
    uint32_t composite;
    int words;
    int o_rel_rows, o_rel_fksel, o_adj, o_inc, o_ec_relids, o_ec_moff, o_ec_soff, o_ec_hasfk;
    int o_mem_fks, o_mem_nd, o_mem_inv_nd, o_mem_skew, o_mem_mcv, o_mem_flags, o_mem_idx_rows,
        o_mem_idx_startup, o_mem_idx_total, o_sels, o_qualtab, o_base;
    /* every class has two members (View::pairs): three 16-byte records per
     * class (PairRec), 4-word aligned */
    int pairs, o_pair_rec;
    /* Tree graphs whose every vertex v > 0 has exactly one lower-numbered
     * neighbour (its parent; true after BFS relabelling of a tree): o_desc
     * holds each vertex's descendants (itself included), o_lglob the vertices
     * of the subtrees of smaller siblings of v and of its ancestors. */
    int is_tree, o_desc, o_lglob;
    /* Trees whose every class of two or more members is one tree edge: the
     * classes crossing the split at edge (parent(c), c) are exactly those on
     * that edge; o_ecm holds their mask (ecw 64-bit words) per vertex c. */
    int tree_ecm, o_ecm;
    /* Tree graphs with at most 4 classes per edge: 8 words per vertex c
     * describing the split at edge (parent(c), c) in both orientations:
     * [0] selectivity (o = 0: c on the inner side), [1] selectivity (o = 1),
     * [2] n_quals | n_classes << 8, [4..5] inner member slots for o = 0,
     * [6..7] for o = 1 (u16 each, class order). */
    int tree_fast, o_tedge;
    /* Rank memo of a tree graph (GQ_MEMO_RANK; 0: none): n words off[t],
     * then n x n words Q[t][u] (attach_rank_tables). */
    int o_rank;
};

struct DPInput {
    BlobHdr hdr;
    std::vector<uint32_t> blob;
    /* base relation values (memo level 1) */
    std::vector<Rel> base;
    /* DPsize: order of the level-1 list (relation indices) */
    std::vector<int> level1;
    /* tree graphs: number of connected sets of each size k (index k = 0..n),
     * from the subtree-count polynomial; empty for other graphs */
    std::vector<uint64_t> tree_counts;
    /* memo layout chosen for this DP: GQ_MEMO_DENSE, or GQ_MEMO_HASH /
     * GQ_MEMO_RANK (trees) */
    int memo = 1;
    /* plain DPsub on a sparse cyclic graph: the splits of each set are
     * enumerated through its biconnected blocks, as MPDP does, and ranked by
     * DPsub's j (dpsub_by_blocks) */
    bool dpsub_blocks = false;
};


/* GQ_MEMO_AUTO resolves to the rank memo for tree graphs whose dense memo
 * (36 B x 2^n) would exceed max_memo_mb; GQ_MEMO_HASH and GQ_MEMO_RANK
 * apply to tree graphs only (other graphs keep the dense memo). */
int resolve_memo(const DPInput &in, int requested, int max_memo_mb);
/* Sets in.memo = resolve_memo(...) and, for the rank memo, appends the rank
 * tables to the blob (BlobHdr::o_rank). */
void prepare_memo(DPInput &in, int requested, int max_memo_mb);

/* Launch-shape knobs of the device DP (gq_tuning in the C ABI); none of
 * them changes a plan, a cost or a count. */
struct Tuning {
    uint32_t lpe2_waves = 4096;    /* tree levels of at most this many two-lane waves take the two-lane form */
    uint32_t gen_grid_max = 512;   /* workgroups of a general-job batch launch (they loop over its tasks) */
    uint32_t tree_iters = 16;      /* wave-iterations of sets per task of the batch tree kernel, and per wave of
                                      the single-query one-lane tree form */
    uint32_t expand_max_n = 20;    /* largest dense tree DP whose level lists are built by expansion */
    uint32_t one_xcd_wgs = 0;      /* single-query tree levels of at most this many workgroups run on one XCD */
};
Tuning tuning_get();
int tuning_set(const Tuning &t);   /* GQ_ERR_INVALID outside the accepted ranges */

struct DPNode {                 /* plan node in DP numbering, pre-order */
    uint32_t set;
    float rows, startup, total;
    int32_t width, left, right, pad;
};

struct DPStats {
    uint64_t candidates = 0, sets = 0, pairs = 0, memo_bytes = 0;
    double gpu_ms = 0.0, filter_ms = 0.0, eval_ms = 0.0;
    int filter_launches = 0, eval_launches = 0;
    int memo_kind = 1;
    /* host clock (steady_clock, ms): the DP's first upload or launch is
     * issued, its last launch is issued, and its results are on the host */
    double t_first_launch = 0.0, t_last_launch = 0.0, t_synced = 0.0;
};

/* steady_clock in milliseconds (host phase stamps of DPStats) */
double host_now_ms();

enum DPAlgo { DP_DPSUB = 0, DP_BICC = 1, DP_DPSIZE = 2 };

/* Split enumeration of a DP: 0 every subset of a set (DPsub), 1 the blocks
 * of a set with MPDP ranks, 2 the blocks with DPsub ranks (the same valid
 * splits as 0, so the same plan). */
inline int dp_mode(const DPInput &in, DPAlgo algo) {
    if (algo == DP_BICC) return 1;
    return (algo == DP_DPSUB && in.dpsub_blocks && !in.hdr.is_tree) ? 2 : 0;
}

/* Run one DP on the current device.  `n_iters` < n_rels selects the IDP1
 * level cap (the tree returned is the cheapest set of that size).  With
 * `noalloc` the DP runs in the library's static scratch array and allocates
 * nothing (dense-memo DPsub/MPDP of at most GQ_NOALLOC_MAX_RELS relations,
 * no IDP1 cap; the device context must already be initialised).  Returns a
 * gq_status. */
int run_dp(const DPInput &in, DPAlgo algo, int n_iters, int memo_kind, bool profile, std::vector<DPNode> &tree,
           DPStats &stats, bool noalloc = false);

/* Run several DPs in one fused launch (one workgroup per DP).  Every input
 * must satisfy fused_eligible(). */
bool fused_eligible(const DPInput &in, DPAlgo algo);

/* Step engine (continuous batching).  Admitted DPs live in slices of one
 * device arena; every step advances each active DP by one level, all of
 * them in one launch.  engine_step enqueues the next step and returns the
 * DPs that finished in the step before it (so the host prepares one step
 * while the device runs the previous one); with no DP active it drains the
 * last step.  DPs join and leave between steps.  Only DPs that satisfy
 * fused_eligible() and run to their full size are admitted.  Not
 * thread-safe: one engine thread drives it. */
struct StepEngine;
struct EngineDone {
    uint64_t key;
    int rc;
    std::vector<DPNode> tree;
    DPStats st;
};
int engine_create(int device, uint64_t arena_bytes, StepEngine **out);
void engine_destroy(StepEngine *e);
/* arena bytes one DP of n relations takes (blob allowance included): dense
 * memo, or open-addressing / rank memo over `total_sets` connected sets */
uint64_t engine_slice_bytes(int n_rels);
uint64_t engine_hash_slice_bytes(uint64_t total_sets);
uint64_t engine_rank_slice_bytes(uint64_t total_sets);
/* admits the DP if its slice fits now (no device call: the blob is copied
 * to the host side of the engine and reaches the device with the next
 * step); false otherwise */
bool engine_admit(StepEngine *e, const DPInput &in, DPAlgo algo, uint64_t key);
int engine_step(StepEngine *e, std::vector<EngineDone> &done);
/* DPs admitted or on the device, plus steps not yet collected: engine_step
 * has work while this is nonzero */
int engine_active(const StepEngine *e);

/* Where run_many sends its GPU DPs when set on the calling thread (the
 * planning server's request threads); unset: the library's own paths. */
struct DPExecutor {
    virtual ~DPExecutor() {}
    virtual int run(const std::vector<const DPInput *> &ins, const std::vector<DPAlgo> &algos,
                    std::vector<std::vector<DPNode>> &trees, std::vector<DPStats> &sts) = 0;
};
DPExecutor *&thread_executor();
int run_dp_batch(const std::vector<const DPInput *> &ins, const std::vector<DPAlgo> &algos,
                 std::vector<std::vector<DPNode>> &trees, std::vector<DPStats> &stats, double *gpu_ms);

}  // namespace gq

#endif
