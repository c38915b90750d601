/*
 * gpuqo_amd.h -- C ABI of the MI355X join-order optimizer core.
 *
 * Plain C types only: fixed-width integers, floats and caller-owned
 * pointers.  Every entry point returns a gq_status; on failure the message is
 * available from gq_last_error() on the calling thread.
 *
 * Relation sets cross the boundary as arrays of 64-bit words, bit i of word
 * (i / 64) standing for relation i (0-based).  The reference encodes relation
 * i as bit i+1 of a Postgres Bitmapset (src/backend/optimizer/gpuqo/
 * gpuqo_bfs_indexing.cu:33); shifting by one bit preserves every numeric
 * order the enumeration depends on, so results are unchanged.
 *
 * Reference interface replaced by this header:
 *   extern "C" QueryTreeC* gpuqo_run(int gpuqo_algorithm, GpuqoPlannerInfoC*)
 *     src/include/optimizer/gpuqo.h:23,
 *     src/backend/optimizer/gpuqo/gpuqo_main_internal.cu:177-185.
 * Input types mirror src/include/optimizer/gpuqo_planner_info.h:20-55 and
 * src/include/optimizer/gpuqo_common.h:91-113; the GUC globals of
 * src/include/optimizer/gpuqo_common.h:38-60 become fields of gq_config.
 */
#ifndef GPUQO_AMD_H
#define GPUQO_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GQ_MAX_RELS 256          /* UnionDP / IDP drivers (host side)          */
#define GQ_MAX_WORDS 4           /* GQ_MAX_RELS / 64                           */
#define GQ_MAX_DP_RELS 32        /* largest single DP handed to the GPU kernels */
#define GQ_NOALLOC_MAX_RELS 20   /* largest query gq_plan_noalloc plans        */

typedef enum gq_status {
    GQ_OK = 0,
    GQ_ERR_INVALID = 1,      /* malformed query or config                  */
    GQ_ERR_UNSUPPORTED = 2,  /* algorithm / size not supported             */
    GQ_ERR_NOMEM = 3,        /* device or host allocation failed           */
    GQ_ERR_DEVICE = 4,       /* HIP runtime error                          */
    GQ_ERR_CAPACITY = 5,     /* output buffer too small                    */
    GQ_ERR_NO_PLAN = 6,      /* query graph disconnected: no join tree     */
    GQ_ERR_INTERRUPTED = 7,  /* interrupt flag raised (CHECK_FOR_INTERRUPTS) */
    GQ_ERR_BUSY = 8,         /* server queue full                          */
    GQ_ERR_INTERNAL = 9      /* unexpected internal failure (e.g. a thread
                                could not be started); no C++ exception
                                ever crosses this interface              */
} gq_status;

/* gpuqo_algorithm values: src/include/optimizer/gpuqo_common.h:13-29. */
typedef enum gq_algorithm {
    GQ_ALGO_DPSIZE = 0,
    GQ_ALGO_DPSUB = 1,
    GQ_ALGO_CPU_DPSIZE = 2,
    GQ_ALGO_CPU_DPSUB = 3,
    GQ_ALGO_CPU_DPSUB_PARALLEL = 4,
    GQ_ALGO_CPU_DPSUB_BICC = 5,
    GQ_ALGO_CPU_DPSUB_BICC_PARALLEL = 6,
    GQ_ALGO_CPU_DPCCP = 7,
    GQ_ALGO_DPE_DPSIZE = 8,
    GQ_ALGO_DPE_DPSUB = 9,
    GQ_ALGO_DPE_DPCCP = 10,
    GQ_ALGO_CPU_GOO = 11
} gq_algorithm;

/* gpuqo_idp_type values: src/include/optimizer/gpuqo_common.h:31-36. */
typedef enum gq_idp_type {
    GQ_IDP1 = 0,
    GQ_IDP2 = 1,
    GQ_UNION = 2
} gq_idp_type;

typedef enum gq_memo_kind {
    GQ_MEMO_AUTO = 0,   /* dense when 2^n slots fit the budget, else rank  */
    GQ_MEMO_DENSE = 1,  /* perfect-hash table indexed by the relation set  */
    GQ_MEMO_HASH = 2,   /* open-addressing table, linear probing           */
    GQ_MEMO_RANK = 3    /* tree graphs: one slot per connected set, the
                           slot computed from the set (no keys, no probes) */
} gq_memo_kind;

/* The join cost function: the reference's make variable cost_function
 * (src/backend/optimizer/gpuqo/Makefile:67-72, gpuqo_cost.cuh:25-38), which
 * selects GPUQO_COST_FUNCTION at build time; here a config field.  SIMPLE
 * is gpuqo_cost_simple.cuh (base relation 0.2 x tuples, nested loop
 * outer.total + outer.rows x inner.total, hash join joinrows + inner.total
 * + outer.total), COUT is gpuqo_cost_cout.cuh (base relation 0, nested loop
 * +inf, hash join as SIMPLE); both have zero startup costs. */
typedef enum gq_cost_function {
    GQ_COST_POSTGRES = 0,
    GQ_COST_SIMPLE = 1,
    GQ_COST_COUT = 2
} gq_cost_function;

/* PathCost, src/include/optimizer/gpuqo_common.h:81-84. */
typedef struct gq_path_cost {
    float startup;
    float total;
} gq_path_cost;

/* BaseRelationC minus the Bitmapset id (relation i is bit i):
 * src/include/optimizer/gpuqo_planner_info.h:31-37. */
typedef struct gq_base_rel {
    float rows;
    float tuples;
    gq_path_cost cost;
    int32_t width;          /* >= 0; the widths of a query sum to at most 2^31-1 (GQ_ERR_INVALID otherwise) */
} gq_base_rel;

/* VarInfo = VarStats + GpuqoIndexInfo:
 * src/include/optimizer/gpuqo_common.h:86-113. */
typedef struct gq_var_info {
    float stadistinct;
    float stanullfrac;
    float mcvfreq;
    uint8_t index_available;
    uint8_t index_unique;
    uint8_t pad_[2];
    float index_rows;
    gq_path_cost index_cost;
} gq_var_info;

/*
 * GpuqoPlannerInfoC (src/include/optimizer/gpuqo_planner_info.h:39-48) in
 * flat arrays.  Equivalence classes are listed in the order the reference
 * iterates them (its EqClassInfo list order).  Members of one class are given
 * as strictly ascending relation indices; per-class arrays are concatenated:
 *   ec_sels : s*(s-1)/2 floats per class, pair order of eqClassIndex
 *             (src/include/optimizer/gpuqo_common.h:65-79)
 *   ec_fks  : one relation set (n_words words) per member
 *   ec_vars : one gq_var_info per member
 */
typedef struct gq_query {
    int32_t n_rels;
    int32_t n_words;                 /* words per relation set = ceil(n/64) */
    const gq_base_rel *base_rels;    /* [n_rels]                            */
    const uint64_t *edge_table;      /* [n_rels * n_words]                  */
    int32_t n_eq_classes;
    const int32_t *ec_member_offsets;/* [n_eq_classes + 1]                  */
    const int32_t *ec_members;       /* [ec_member_offsets[n_eq_classes]]   */
    const float *ec_sels;
    const uint64_t *ec_fks;          /* [n_members * n_words]               */
    const gq_var_info *ec_vars;      /* [n_members]                         */
} gq_query;

/*
 * GUC mirror.  Names and defaults follow src/backend/utils/misc/guc.c
 * (bools :1102-1171, ints :2139-2298, enums :4809-4826) and the Postgres cost
 * GUCs snapshotted by setParams (gpuqo_planner_info.cu:113-131).
 */
typedef struct gq_config {
    int32_t gpuqo_algorithm;                      /* GQ_ALGO_*, default DPSIZE */
    int32_t gpuqo_dpsub_bicc;                     /* off                       */
    int32_t gpuqo_dpsub_filter;                   /* on                        */
    int32_t gpuqo_dpsub_filter_threshold;         /* 16                        */
    int32_t gpuqo_dpsub_filter_cpu_enum_threshold;/* 1024                      */
    int32_t gpuqo_dpsub_filter_keys_overprovisioning; /* 100                   */
    int32_t gpuqo_dpsub_ccc;                      /* on                        */
    int32_t gpuqo_dpsub_csg;                      /* on                        */
    int32_t gpuqo_dpsub_csg_threshold;            /* 128                       */
    int32_t gpuqo_dpsub_tree;                     /* off: accepted; tree sets
                                                     are always costed per edge
                                                     here, in CPU DPsub split
                                                     order                     */
    int32_t gpuqo_spanning_tree;                  /* off: plan over the minimum
                                                     spanning tree of the join
                                                     graph (weight: base-pair
                                                     join rows); costs still
                                                     use every class           */
    int32_t gpuqo_n_parallel;                     /* 10240                     */
    int32_t gpuqo_scratchpad_size_mb;             /* 10                        */
    int32_t gpuqo_min_memo_size_mb;               /* 1                         */
    int32_t gpuqo_max_memo_size_mb;               /* 1024                      */
    int32_t gpuqo_idp_type;                       /* GQ_IDP2                   */
    int32_t gpuqo_idp_n_iters;                    /* 0 = off                   */
    int32_t gpuqo_dpe_n_threads;                  /* 4                         */
    int32_t gpuqo_cpu_dpsub_parallel_chunk_size;  /* 64                        */
    int32_t gpuqo_dpe_pairs_per_depbuf;           /* 128                       */
    /* Postgres cost GUCs (double in Postgres, float once inside gpuqo). */
    double seq_page_cost;                         /* 1.0     */
    double random_page_cost;                      /* 4.0     */
    double cpu_tuple_cost;                        /* 0.01    */
    double cpu_index_tuple_cost;                  /* 0.005   */
    double cpu_operator_cost;                     /* 0.0025  */
    double disable_cost;                          /* 1.0e10  */
    int32_t effective_cache_size;                 /* 524288  */
    int32_t work_mem;                             /* 4096 kB */
    uint8_t enable_seqscan, enable_indexscan, enable_tidscan, enable_sort;
    uint8_t enable_hashagg, enable_nestloop, enable_mergejoin, enable_hashjoin;
    /* Build-specific knobs (no reference counterpart). */
    int32_t memo_kind;                            /* GQ_MEMO_AUTO              */
    int32_t device;                               /* HIP device ordinal, 0     */
    int32_t profile_kernels;                      /* time each phase with HIP
                                                     events (stats.*_ms)       */
    /* The reference's build-time cost function (Makefile cost_function). */
    int32_t cost_function;                        /* GQ_COST_POSTGRES          */
} gq_config;

/* One node of the returned join tree (QueryTreeC,
 * src/include/optimizer/gpuqo_planner_info.h:50-57).  Nodes are stored in
 * pre-order; node 0 is the root; leaves have left = right = -1. */
typedef struct gq_plan_node {
    uint64_t relids[GQ_MAX_WORDS];
    float rows;
    gq_path_cost cost;
    int32_t width;
    int32_t left;
    int32_t right;
} gq_plan_node;

/* Per-call counters and timings (no reference counterpart; the reference
 * prints the same phases under GPUQO_PROFILE, gpuqo_timing.cuh:22-55). */
typedef struct gq_plan_stats {
    uint64_t candidates;      /* C: relation sets unranked and filtered    */
    uint64_t sets;            /* S: connected sets evaluated (size >= 2)   */
    uint64_t pairs;           /* P: ordered valid csg-cmp pairs costed     */
    uint64_t memo_bytes;      /* device bytes held by the memo             */
    double gpu_ms;            /* device time of the DP: HIP events around
                                 its launches, or, for a dense tree DP planned
                                 whole without profile_kernels, the device
                                 clock from its first launch's start to the
                                 end of its plan walk                       */
    double wall_ms;           /* host wall time of the whole call          */
    int32_t n_dp_calls;       /* DPs run (1 for a plain query)             */
    int32_t memo_kind;        /* GQ_MEMO_DENSE, _HASH or _RANK used        */
    /* with cfg->profile_kernels: summed device time per phase */
    double filter_ms;         /* unrank + connectivity filter kernels      */
    double eval_ms;           /* split evaluation (+ finalize) kernels     */
    int32_t filter_launches;
    int32_t eval_launches;
    /* host phases of wall_ms: from the call's start until the first DP's
     * first upload is issued; from then until the last DP's last launch is
     * issued; from the last DP's results arriving until the call returns */
    double host_prep_ms;
    double host_enqueue_ms;
    double host_post_ms;
} gq_plan_stats;

typedef struct gq_plan_out {
    gq_plan_node *nodes;      /* caller-owned, capacity >= 2*n_rels-1      */
    int32_t capacity;
    int32_t n_nodes;          /* written                                   */
    gq_plan_stats stats;      /* written                                   */
} gq_plan_out;

/* Fill *cfg with the reference GUC defaults (guc.c, see gq_config). */
gq_status gq_config_default(gq_config *cfg);

/* Library / device set-up; idempotent.  Called implicitly by gq_plan. */
gq_status gq_init(int32_t device);

/*
 * Plan one query.  Replaces gpuqo_run (gpuqo_main_internal.cu:177-185):
 * converts the input, BFS-relabels from relation 0, routes to the DP or to
 * IDP1 / IDP2 / UnionDP as gpuqo_idp_n_iters and gpuqo_idp_type say
 * (gpuqo_main_internal.cu:81-147), and returns the join tree in the caller's
 * numbering.  GQ_ALGO_DPSUB and GQ_ALGO_DPSIZE run on the GPU.
 */
gq_status gq_plan(const gq_query *q, const gq_config *cfg, gq_plan_out *out);

/*
 * gq_plan without any allocation.  No counterpart in the reference: this
 * entry point is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * plans like gq_plan -- same algorithms, same plans bit for bit -- but the
 * DP's device state lives in a static device array of about 40 MB that every
 * GPU the library runs on holds until the process ends, and the call makes no
 * hipMalloc, hipFree or pinned-memory allocation (the query's conversion on
 * the host uses the host heap only).  Limits: GQ_ALGO_DPSUB (plain or
 * gpuqo_dpsub_bicc) on at most GQ_NOALLOC_MAX_RELS relations, no IDP
 * decomposition, the dense memo; anything else returns GQ_ERR_UNSUPPORTED
 * before any device work.  gq_init(cfg->device) must have been called
 * (GQ_ERR_INVALID otherwise).  Calls on one GPU share its array, which keeps
 * the last call's contents; they are serialised by the library's device
 * lock, so overlapping calls wait rather than race.
 */
gq_status gq_plan_noalloc(const gq_query *q, const gq_config *cfg, gq_plan_out *out);

/*
 * Plan n_queries independent queries in one batched device pass: all DPs of
 * the batch advance level by level together, one launch per level bucket.
 * Each result equals what gq_plan returns for that query alone.  No
 * reference counterpart; per-query semantics are those of gpuqo_run.
 */
gq_status gq_plan_batch(int32_t n_queries, const gq_query *const *qs,
                        const gq_config *cfg, gq_plan_out *outs);

/*
 * Continuous-batching planning server (no reference counterpart).  Requests
 * are admitted by projected memo bytes, run concurrently on the device, and
 * complete out of order.  A request's plan equals gq_plan's for the same
 * query.  Queries are copied at submit time.
 */
typedef struct gq_server gq_server;
gq_status gq_server_create(const gq_config *cfg, uint64_t memo_budget_bytes,
                           int32_t max_inflight, gq_server **out);
gq_status gq_server_submit(gq_server *s, const gq_query *q, uint64_t tag);
/* Submit with a per-request configuration (NULL: the server's); its device
 * must be the server's.  Lets one server mix plain DPs and UnionDP/IDP2. */
gq_status gq_server_submit_cfg(gq_server *s, const gq_query *q, const gq_config *cfg, uint64_t tag);
/* Poll one finished request; returns GQ_OK with *done=1 when one was
 * written, *done=0 when none is ready.  wait: 0 returns at once, > 0 waits
 * up to that many milliseconds, < 0 waits while requests are outstanding.
 * out->nodes must hold 2n-1 nodes of the largest query submitted (else
 * GQ_ERR_CAPACITY, and the request stays queued).  out->stats.wall_ms is the
 * request's latency from submission to completion. */
gq_status gq_server_poll(gq_server *s, int32_t wait, int32_t *done,
                         uint64_t *tag, gq_status *req_status,
                         gq_plan_out *out);
gq_status gq_server_destroy(gq_server *s);

/*
 * Memo snapshots.  They replace the reference's per-node HashTable::get
 * walk of the memo (gpuqo_hashtable.cu:148-178, dpsub_buildQueryTree in
 * gpuqo_query_tree.cuh:46-72) with a device-resident copy of the whole
 * memo: every connected set of the query (the base relations included),
 * with its best join -- outer side, rows, startup and total cost, width --
 * in the caller's relation numbering, kept in device memory until
 * gq_snapshot_free.  The plan of any connected sub-query (the join of a
 * connected subset of the relations) is then walked from it on the device
 * without planning again: that entry is the DP's optimum for those
 * relations alone.  A snapshot can be written out in the format below and
 * read back, on this process or another.
 *
 * gq_plan_snapshot: gq_plan (the same plan, bit for bit, in *out) plus the
 * snapshot of its DP.  One DPsub or MPDP DP of 2..64 relations: no IDP
 * (gpuqo_idp_n_iters), no DPsize (GQ_ERR_UNSUPPORTED).
 * gq_snapshot_subplan: the plan of the connected sub-query `relids` (one
 * word, GQ_MAX_WORDS read as gq_query's; bits outside the query's
 * relations or a set that is not connected give GQ_ERR_INVALID); stats
 * are zero.
 * gq_snapshot_export: *bytes = the export's size; with buf and cap >=
 * *bytes, writes it (GQ_ERR_CAPACITY when cap is smaller).
 * gq_snapshot_import: a snapshot from an export (GQ_ERR_INVALID when the
 * buffer is not one).
 *
 * Export format, little-endian:
 *   bytes 0..7   magic "GQMEMO1" and a zero byte;
 *   8..11        version (1);       12..15  relations n (1..64);
 *   16..23       records m;         24..27  bytes per record (32);
 *   28..63       zero;
 *   then m records of 32 bytes, ordered by relation count, then by set:
 *     u64 set (bit i = the caller's relation i), u64 outer side of the
 *     set's best join (0 for a base relation), f32 rows, f32 startup cost,
 *     f32 total cost, i32 width.
 */
typedef struct gq_snapshot gq_snapshot;
#define GQ_SNAPSHOT_MAGIC "GQMEMO1"
#define GQ_SNAPSHOT_VERSION 1u
#define GQ_SNAPSHOT_HEADER_BYTES 64u
#define GQ_SNAPSHOT_RECORD_BYTES 32u
gq_status gq_plan_snapshot(const gq_query *q, const gq_config *cfg, gq_plan_out *out, gq_snapshot **snap);
gq_status gq_snapshot_subplan(const gq_snapshot *snap, const uint64_t *relids, gq_plan_out *out);
gq_status gq_snapshot_info(const gq_snapshot *snap, int32_t *n_rels, uint64_t *n_sets, uint64_t *device_bytes);
gq_status gq_snapshot_export(const gq_snapshot *snap, void *buf, uint64_t cap, uint64_t *bytes);
gq_status gq_snapshot_import(const void *buf, uint64_t bytes, gq_snapshot **snap);
gq_status gq_snapshot_free(gq_snapshot *snap);

/*
 * Launch-shape knobs of the device DP, per process.  None of them changes a
 * plan, a cost or a count.  They take the place of the reference's
 * hand-set gpuqo_n_parallel / gpuqo_scratchpad_size_mb /
 * gpuqo_dpsub_filter_keys_overprovisioning GUCs (guc.c:2150-2217), which
 * size its own kernels and have no meaning here.
 *   tree_two_lane_waves   tree levels of at most this many waves at two
 *                         lanes per split take the two-lane form (1..2^24)
 *   general_grid_max      workgroups of a general-graph batch launch (1..65536)
 *   tree_task_iters       wave-iterations of sets per batch tree task (1..256)
 *   tree_expand_max_rels  largest dense tree DP whose level lists are built
 *                         by expansion instead of the 2^n filter (0..32)
 *   tree_one_xcd_wgs      single-query tree levels of at most this many
 *                         workgroups run on one XCD (8x the grid, every
 *                         eighth workgroup working), so a level's writes
 *                         stay in the L2 the next level reads (0..256; 0 off)
 */
typedef struct gq_tuning {
    int32_t tree_two_lane_waves;
    int32_t general_grid_max;
    int32_t tree_task_iters;
    int32_t tree_expand_max_rels;
    int32_t tree_one_xcd_wgs;
} gq_tuning;
gq_status gq_tuning_get(gq_tuning *t);
gq_status gq_tuning_set(const gq_tuning *t);
/*
 * Autotuner (no reference counterpart): times the sample queries under
 * candidate knob values, one knob at a time, keeps the fastest value of
 * each (a knob leaves its current value only for one at least 1% faster),
 * leaves it set and returns all knobs in *out.  `knobs` selects the groups:
 * GQ_TUNE_SINGLE (tree_two_lane_waves, tree_expand_max_rels,
 * tree_one_xcd_wgs; timed by the device time of planning the queries one by
 * one with gq_plan, 4 * reps timings) and GQ_TUNE_BATCH (tree_task_iters,
 * general_grid_max; timed by the wall time of one gq_plan_batch of them,
 * reps timings).  The median counts.  Must not run concurrently with other
 * planning calls of the process.
 */
enum { GQ_TUNE_SINGLE = 1, GQ_TUNE_BATCH = 2 };
gq_status gq_autotune(int32_t n_queries, const gq_query *const *qs, const gq_config *cfg, int32_t reps,
                      int32_t knobs, gq_tuning *out);

/* Request an interrupt of the running call (CHECK_FOR_INTERRUPTS,
 * gpuqo.cuh:21-28): the planner stops enqueuing DP levels (the levels
 * already enqueued finish), and the call returns GQ_ERR_INTERRUPTED; the
 * device buffers stay cached for the next call.  gq_interrupt(0) clears the
 * flag. */
void gq_interrupt(int32_t on);

const char *gq_last_error(void);
const char *gq_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GPUQO_AMD_H */
