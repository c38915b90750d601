/*
 * gq_wide.h -- host-side interface of the wide DP: DPsub / MPDP over one
 * query of 33 to 64 relations on 64-bit relation sets (gq_wide.hip).
 * Internal to the library; the public C ABI is include/gpuqo_amd.h.
 *
 * The reference runs its GPU DPsub over Bitmapset64 for queries of up to 63
 * relations (gpuqo_main_internal.cu:178-184), unranking every k-subset of
 * every level (gpuqo_dpsub.cuh:95-126) and filtering the connected ones.
 * Here the level lists hold only the connected sets; which queries that is
 * practical for is set by their number of connected sets.  Tree join
 * graphs build each set once from a set of the level below, with level
 * sizes the host counts exactly; other graphs (cycles, chords) build each
 * set from every set of the level below that it extends, keep the copy
 * whose memo insert comes first, and learn each level's size after it is
 * built.
 */
#ifndef GQ_WIDE_H
#define GQ_WIDE_H

#include <stdint.h>

#include <functional>
#include <vector>

#include "gq_dp.h"

namespace gq {

template <int W>
struct HQuery;

#define GQ_WIDE_MAX_RELS 64

/* Query description of a wide DP (32-bit words; 64-bit fields 2-word
 * aligned, the PairRec records 4-word aligned), staged into LDS.  The
 * arrays are those of View<uint64_t> (gq_cost_model.h), plus:
 *   o_base   base relation i: rows, startup, total, width (4 words);
 *   o_desc   per vertex v, v and its descendants (u64) -- the parent of
 *            v > 0 is its one lower-numbered neighbour;
 *   o_lglob  per vertex v, the vertices of the subtrees of the smaller
 *            siblings of v and of each of its ancestors (u64);
 *   o_kids   per vertex v, its children (u64);
 *   o_adj    per vertex v, its neighbours (u64; every graph);
 *   o_tedge  (tree_fast) per vertex c, the split at edge (parent(c), c):
 *            [0] selectivity with c's side inner, [1] with it outer,
 *            [2] n_quals | n_members << 8, [4..5] / [6..7] the inner member
 *            slots of the two orientations (u16, class order). */
struct WideHdr {
    int n_rels, n_ec, n_mem, ecw, words;
    Params p;
    uint64_t composite;
    int o_rel_rows, o_rel_fksel, o_base, o_inc, o_ec_relids, o_ec_moff, o_ec_soff, o_ec_hasfk;
    int o_mem_fks, o_mem_nd, o_mem_inv_nd, o_mem_skew, o_mem_mcv, o_mem_flags, o_mem_idx_rows,
        o_mem_idx_startup, o_mem_idx_total, o_sels, o_qualtab;
    int pairs, o_pair_rec;
    int o_desc, o_lglob, o_kids;
    int tree_fast, o_tedge;
    int is_tree;    /* a tree numbered parents first: the tables above are set */
    int o_adj;
};

struct WideInput {
    WideHdr hdr;
    std::vector<uint32_t> blob;
    /* trees: connected sets of each size k = 0..n (subtree-count
     * polynomial); empty for other graphs */
    std::vector<uint64_t> counts;
};

/* plan node in DP numbering; node 0 is the root, children follow their
 * parent (breadth-first) */
struct WideNode {
    uint64_t set;
    float rows, startup, total;
    int32_t width, left, right;
};

/* A relabelled query (BFS order from relation 0) of 2..64 relations: a
 * tree numbered parents first, or any connected graph (the general path);
 * GQ_ERR_UNSUPPORTED above the class limits. */
int build_wide_input(const HQuery<1> &q, WideInput &in);

/* Device bytes the DP takes: the open-addressing memo (32 B slots, load at
 * most 1/2) and the level lists. */
uint64_t wide_dp_bytes(const WideInput &in);

/* Runs the DP (DPsub or MPDP order) on the library's device stream;
 * GQ_ERR_NOMEM when wide_dp_bytes exceeds max_memo_mb. */
int run_wide_dp(const WideInput &in, DPAlgo algo, int max_memo_mb, std::vector<WideNode> &nodes, DPStats &st);

/* ---------------------------------------------------------- memo snapshots
 *
 * A finished DP's memo copied into an open-addressing table of WSlot
 * records (gq_device.h) keyed by every connected set of the query in the
 * caller's relation numbering, with the outer side of the set's chosen
 * split in the same numbering (0 for base relations): the plan of every
 * connected sub-query can be walked from it without planning again. */
struct SnapTable {
    void *tab = nullptr;      /* WSlot[2^lg] on `device` */
    uint32_t lg = 0;
    uint64_t n_sets = 0;      /* connected sets, the base relations included */
    int n_rels = 0, device = -1;
};

/* Set by gq_plan_snapshot around the one DP of its call: DP relation i is
 * the caller's relation perm[i].  run_dp and run_wide_dp, when the DP they
 * finish has n relations, build `out` on the device before they return
 * (and set done). */
struct SnapReq {
    int n = 0;
    int8_t perm[64];
    SnapTable out;
    bool done = false;
};
extern thread_local SnapReq *g_snap_req;

/* Slots for n_sets sets at a load of at most 1/2. */
uint32_t snap_lg(uint64_t n_sets);

/* Plan of the connected sub-query `root` (caller's numbering) walked from
 * the snapshot, breadth-first (WideNode); GQ_ERR_INVALID if root is not one
 * of its sets. */
int snapshot_walk(const SnapTable &t, uint64_t root, std::vector<WideNode> &nodes);

/* Every record of the snapshot, sorted by (size, set): set, left, rows,
 * startup, total, width. */
struct SnapRecord {
    uint64_t set, left;
    float rows, startup, total;
    int32_t width;
};
int snapshot_read(const SnapTable &t, std::vector<SnapRecord> &recs);

/* A snapshot on the device from records (an imported one). */
int snapshot_build(const std::vector<SnapRecord> &recs, int n_rels, SnapTable &t);

void snapshot_free(SnapTable &t);

/* f(device, stream) under the device context's lock, on its stream
 * (gq_dp.hip) */
int with_device_stream(const std::function<int(int, void *)> &f);

}  // namespace gq

#endif
