/*
 * gq_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's join-order search: the expected plans
 * in the GPU product's parity tests and the timed CPU baseline.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (postgres-gpuqo_amd/) never links or calls it.
 *
 * The reference (Manciukic/postgres-gpuqo, src/backend/optimizer/gpuqo/, below
 * "gpuqo/") cannot be compiled in this image: its headers include the
 * configure-generated pg_config.h and Postgres runtime symbols.  Everything
 * here is therefore a from-scratch restatement; each function cites the
 * reference lines it follows.  Relation i is bit i+1 of a set, as in the
 * reference (gpuqo/gpuqo_bfs_indexing.cu:33).
 *
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no -march=native):
 * the float cost model must not be contracted into FMAs.
 */
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cfloat>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <queue>
#include <stack>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/gpuqo_amd.h"

namespace oracle {

/* ------------------------------------------------------------------------
 * Relation sets: W 64-bit words, bit i+1 = relation i.
 * Operations follow gpuqo/gpuqo_bitmapset.cuh:21-310.
 * ---------------------------------------------------------------------- */
template <int W>
struct Bits {
    uint64_t w[W];
    static constexpr int SIZE = 64 * W;

    Bits() { for (int i = 0; i < W; i++) w[i] = 0; }
    explicit Bits(uint64_t v) { w[0] = v; for (int i = 1; i < W; i++) w[i] = 0; }

    static Bits nth(int n) { Bits b; b.w[n >> 6] = 1ULL << (n & 63); return b; }
    int size() const { int s = 0; for (int i = 0; i < W; i++) s += __builtin_popcountll(w[i]); return s; }
    bool empty() const { for (int i = 0; i < W; i++) if (w[i]) return false; return true; }
    int lowestPos() const {             /* ffs - 1; -1 when empty */
        for (int i = 0; i < W; i++) if (w[i]) return i * 64 + __builtin_ctzll(w[i]);
        return -1;
    }
    Bits lowest() const { Bits b; int p = lowestPos(); if (p >= 0) b = nth(p); return b; }
    Bits allLower() const {             /* bits strictly below the lowest set bit */
        Bits b; int p = lowestPos();
        for (int i = 0; i < W; i++) {
            if (p >= (i + 1) * 64) b.w[i] = ~0ULL;
            else if (p > i * 64) b.w[i] = (1ULL << (p - i * 64)) - 1;
        }
        return b;
    }
    Bits allLowerInc() const { Bits b = allLower(); int p = lowestPos(); if (p >= 0) b.set(p); return b; }
    Bits operator|(const Bits &o) const { Bits b; for (int i = 0; i < W; i++) b.w[i] = w[i] | o.w[i]; return b; }
    Bits operator&(const Bits &o) const { Bits b; for (int i = 0; i < W; i++) b.w[i] = w[i] & o.w[i]; return b; }
    Bits operator-(const Bits &o) const { Bits b; for (int i = 0; i < W; i++) b.w[i] = w[i] & ~o.w[i]; return b; }
    Bits &operator|=(const Bits &o) { for (int i = 0; i < W; i++) w[i] |= o.w[i]; return *this; }
    Bits &operator&=(const Bits &o) { for (int i = 0; i < W; i++) w[i] &= o.w[i]; return *this; }
    Bits &operator-=(const Bits &o) { for (int i = 0; i < W; i++) w[i] &= ~o.w[i]; return *this; }
    bool operator==(const Bits &o) const { for (int i = 0; i < W; i++) if (w[i] != o.w[i]) return false; return true; }
    bool operator!=(const Bits &o) const { return !(*this == o); }
    bool operator<(const Bits &o) const {
        for (int i = W - 1; i >= 0; i--) if (w[i] != o.w[i]) return w[i] < o.w[i];
        return false;
    }
    bool intersects(const Bits &o) const { for (int i = 0; i < W; i++) if (w[i] & o.w[i]) return true; return false; }
    bool isSubset(const Bits &o) const { return (*this & o) == *this; }   /* this ⊆ o */
    void set(int n) { w[n >> 6] |= 1ULL << (n & 63); }
    void unset(int n) { w[n >> 6] &= ~(1ULL << (n & 63)); }
    bool isSet(int n) const { return (w[n >> 6] >> (n & 63)) & 1ULL; }
    /* multiword a - b (two's complement) */
    static Bits sub(const Bits &a, const Bits &b) {
        Bits r; unsigned borrow = 0;
        for (int i = 0; i < W; i++) {
            uint64_t x = a.w[i], y = b.w[i];
            uint64_t d = x - y - borrow;
            borrow = (x < y) || (x == y && borrow) ? 1 : 0;
            r.w[i] = d;
        }
        return r;
    }
    static Bits add1(const Bits &a) {
        Bits r = a;
        for (int i = 0; i < W; i++) { r.w[i] += 1; if (r.w[i] != 0) break; }
        return r;
    }
};

template <int W>
struct BitsHash {
    size_t operator()(const Bits<W> &b) const {
        /* W == 1: identity, as std::hash<unsigned long long> (and therefore
         * std::hash<GpuqoBitmapset<T>>, gpuqo_bitmapset.cuh:323-330). */
        uint64_t h = b.w[0];
        for (int i = 1; i < W; i++) h = h * 0x9E3779B97F4A7C15ULL ^ b.w[i];
        return (size_t)h;
    }
};

/* nextSubset: gpuqo_bitmapset.cuh:300-304 */
template <int W>
static inline Bits<W> nextSubset(const Bits<W> &subset, const Bits<W> &set) {
    return set & Bits<W>::sub(subset, set);
}

/* Gosper next permutation: gpuqo_bitmapset.cuh:146-149 (single word) */
static inline uint64_t next_perm(uint64_t v) {
    uint64_t t = (v | (v - 1)) + 1;
    return t | ((((t & (~t + 1)) / (v & (~v + 1))) >> 1) - 1);
}

/* pdep loop emulation: gpuqo_bit_manipulation.cuh:98-122 */
static inline uint64_t pdep64(uint64_t a, uint64_t mask) {
    uint64_t res = 0;
    for (uint64_t bb = 1; mask; bb += bb) {
        if (a & bb) res |= mask & (~mask + 1);
        mask &= mask - 1;
    }
    return res;
}
template <int W>
static Bits<W> expandToMask(int idx, const Bits<W> &mask) {  /* nth(idx) deposited */
    Bits<W> m = mask;
    for (int i = 0; i < idx; i++) m -= m.lowest();
    return m.lowest();
}

/* ------------------------------------------------------------------------
 * Planner info: gpuqo_planner_info.cuh:22-137
 * ---------------------------------------------------------------------- */
struct PathCost { float startup, total; };

struct Params {           /* GpuqoPlannerInfoParams, planner_info.cuh:79-96 */
    float effective_cache_size, seq_page_cost, random_page_cost, cpu_tuple_cost,
          cpu_index_tuple_cost, cpu_operator_cost, disable_cost;
    bool enable_seqscan, enable_indexscan, enable_tidscan, enable_sort,
         enable_hashagg, enable_nestloop, enable_mergejoin, enable_hashjoin;
    int work_mem;
    int cost_function;    /* GPUQO_COST_FUNCTION (gpuqo_cost.cuh:25-38): 0 postgres, 1 simple, 2 cout */
};

template <int W>
struct BaseRel {
    Bits<W> id;
    float rows, tuples;
    int width;
    PathCost cost;
    bool composite;
};

template <int W>
struct QueryTree {
    Bits<W> id;
    float rows;
    PathCost cost;
    int width;
    QueryTree *left, *right;
};

template <int W>
static void freeQueryTree(QueryTree<W> *qt) {
    if (!qt) return;
    freeQueryTree(qt->left);
    freeQueryTree(qt->right);
    delete qt;
}

template <int W>
struct Info {
    int n_rels = 0, n_iters = 0;
    Params params;
    std::vector<BaseRel<W>> base_rels;
    std::vector<Bits<W>> edge_table;
    int n_ec = 0;
    std::vector<Bits<W>> ec_relids;
    std::vector<float> ec_sels;
    std::vector<Bits<W>> ec_fks;
    std::vector<gq_var_info> ec_vars;
};

static inline int eqClassNSels(int size) { return size * (size - 1) / 2; }
static inline int eqClassIndex(int l, int r, int size) {   /* gpuqo_common.h:71-79 */
    if (l > r) { int t = l; l = r; r = t; }
    return l * size - l * (l + 1) / 2 + (r - l - 1);
}

/* ------------------------------------------------------------------------
 * Cost model: gpuqo_cost.cuh, gpuqo_cost_postgres.cuh,
 * gpuqo_row_estimation.cuh.  float arithmetic in the reference's order.
 * The reference's host build resolves min/max on float to fminf/fmaxf.
 * ---------------------------------------------------------------------- */
struct JoinRel {                    /* JoinRelation (value part) */
    float rows;
    PathCost cost;
    int width;
};

struct CostExtra { bool inner_unique, indexed_join_quals; float joinrows, outer_match_frac, match_count; };
struct QualCost { int n_quals; float startup, per_tuple; };
struct BucketStats { float mcvfreq, bucketsize; };
struct HtszRes { int numbuckets, numbatches, num_skew_mcvs; };

static const float NANF_ = __builtin_nanf("");
static const float INFF_ = __builtin_inff();

#define MAXALIGN8(len) ((((uintptr_t)(len)) + 7) & ~((uintptr_t)7))
static const long MAX_ALLOC_SIZE = 0x3fffffffL;
static const int BLCKSZ_ = 8192;

/* x86 cvttss2si: out-of-range or NaN gives INT_MIN.  The reference passes a
 * float to an int parameter at cost_postgres.cuh:370 / :624-626. */
static inline int float_to_int_x86(float f) {
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int)f;
    return INT_MIN;
}

static inline int floor_log2(long num) { return 63 - __builtin_clzll((unsigned long long)num); }  /* :89-93 */
static inline int ceil_log2(long num) {                                                          /* :95-101 */
    int res = floor_log2(num);
    return (1L << res) == num ? res : res + 1;
}

/* clamp_row_est: cost_postgres.cuh:850-863 */
static float clamp_row_est(float nrows) {
    if (nrows <= 1.0f) nrows = 1.0f;
    else nrows = rintf(nrows);
    return nrows;
}
/* relation_byte_size / page_size: :822-840 */
static float relation_byte_size(float tuples, int width) {
    return tuples * (float)(MAXALIGN8(width) + 24);
}
static float page_size(float tuples, int width) {
    return ceilf(relation_byte_size(tuples, width) / (float)BLCKSZ_);
}

/* ExecChooseHashTableSize: cost_postgres.cuh:479-618 */
static HtszRes exec_choose_hash_table_size(float ntuples, int tupwidth, int work_mem) {
    HtszRes res;
    res.numbatches = 1;
    if (ntuples <= 0.0f) ntuples = 1000.0f;
    int tupsize = (int)(32 + MAXALIGN8(tupwidth));
    float inner_rel_bytes = ntuples * tupsize;
    long hash_table_bytes = work_mem * 1024L;
    float space_allowed = (float)hash_table_bytes;
    long skew_table_bytes = hash_table_bytes * 2 / 100;
    res.num_skew_mcvs = (int)(skew_table_bytes / (tupsize + 84));
    if (res.num_skew_mcvs > 0) hash_table_bytes -= skew_table_bytes;
    long max_pointers = (long)(space_allowed / (float)sizeof(void *));
    max_pointers = std::min(max_pointers, (long)(MAX_ALLOC_SIZE / sizeof(void *)));
    long mppow2 = 1L << ceil_log2(max_pointers);
    if (max_pointers != mppow2) max_pointers = mppow2 / 2;
    max_pointers = std::min(max_pointers, (long)INT_MAX / 2);
    float dbuckets = ceilf(ntuples / 1);
    dbuckets = fminf(dbuckets, (float)max_pointers);
    res.numbuckets = float_to_int_x86(dbuckets);
    res.numbuckets = std::max(res.numbuckets, 1024);
    res.numbuckets = 1 << ceil_log2(res.numbuckets);
    long bucket_bytes = (long)(sizeof(void *) * (size_t)(long)res.numbuckets);
    if (inner_rel_bytes + (float)bucket_bytes > (float)hash_table_bytes) {
        long bucket_size = (long)(tupsize * 1 + sizeof(void *));
        long lbuckets = 1L << ceil_log2(hash_table_bytes / bucket_size);
        lbuckets = std::min(lbuckets, max_pointers);
        res.numbuckets = (int)lbuckets;
        res.numbuckets = 1 << ceil_log2(res.numbuckets);
        bucket_bytes = (long)((size_t)(long)res.numbuckets * sizeof(void *));
        float dbatch = ceilf(inner_rel_bytes / (float)(hash_table_bytes - bucket_bytes));
        dbatch = fminf(dbatch, (float)max_pointers);
        int minbatch = float_to_int_x86(dbatch);
        res.numbatches = std::max(2, 1 << ceil_log2(minbatch));
    }
    return res;
}

/* __estimate_hash_bucketsize: cost_postgres.cuh:701-775 */
template <int W>
static float estimate_hash_bucketsize(const gq_var_info &v, const BaseRel<W> &baserel, int nbuckets) {
    float estfract, ndistinct, mcvfreq, avgfreq;
    ndistinct = v.stadistinct;
    if (ndistinct < 0.0f) ndistinct = -ndistinct * baserel.tuples;
    if (ndistinct <= 0.0f) return 0.1f;
    avgfreq = (1.0f - v.stanullfrac) / ndistinct;
    if (baserel.tuples > 0.0f) {
        ndistinct *= baserel.rows / baserel.tuples;
        ndistinct = clamp_row_est(ndistinct);
    }
    if (ndistinct > (float)nbuckets) estfract = 1.0f / (float)nbuckets;
    else estfract = 1.0f / ndistinct;
    mcvfreq = v.mcvfreq;
    if (avgfreq > 0.0f && mcvfreq > avgfreq) estfract *= mcvfreq / avgfreq;
    if (estfract < 1.0e-6f) estfract = 1.0e-6f;
    else if (estfract > 1.0f) estfract = 1.0f;
    return estfract;
}

/* index of the lowest relation of `match` among the members of `ec` */
template <int W>
static inline int member_idx(const Bits<W> &match, const Bits<W> &ec) {
    return (match.lowest().allLower() & ec).size();
}

/* estimate_hash_bucket_stats: cost_postgres.cuh:620-668 */
template <int W>
static BucketStats estimate_hash_bucket_stats(const Bits<W> &outer, const Bits<W> &inner, int nbuckets,
                                              const Info<W> &info) {
    BucketStats st; st.bucketsize = 1.0f; st.mcvfreq = 1.0f;
    for (int i = 0, off_vars = 0; i < info.n_ec; off_vars += info.ec_relids[i].size(), i++) {
        const Bits<W> &ec = info.ec_relids[i];
        Bits<W> ml = ec & outer, mr = ec & inner;
        if (ml.empty() || mr.empty()) continue;
        while (!mr.empty()) {
            Bits<W> in_id = mr.lowest();
            int in_idx = (in_id.allLower() & ec).size();
            const BaseRel<W> &br = info.base_rels[in_id.lowestPos() - 1];
            const gq_var_info &var = info.ec_vars[off_vars + in_idx];
            float t = estimate_hash_bucketsize(var, br, nbuckets);
            if (st.bucketsize > t) st.bucketsize = t;
            if (st.mcvfreq > var.mcvfreq) st.mcvfreq = var.mcvfreq;
            mr -= in_id;
        }
    }
    return st;
}

/* cost_qual_eval: cost_postgres.cuh:786-815 */
template <int W>
static QualCost cost_qual_eval(const Bits<W> &l, const Bits<W> &r, const Info<W> &info) {
    QualCost c; c.startup = 0.0f; c.per_tuple = 0.0f; c.n_quals = 0;
    for (int i = 0; i < info.n_ec; i++) {
        const Bits<W> &ec = info.ec_relids[i];
        if ((ec & l).empty() || (ec & r).empty()) continue;
        c.per_tuple += info.params.cpu_operator_cost;
        c.n_quals++;
    }
    return c;
}

/* cost_nestloop: cost_postgres.cuh:110-275 */
template <int W>
static PathCost cost_nestloop(const Bits<W> &outer_id, const JoinRel &outer, const Bits<W> &inner_id,
                              const JoinRel &inner, const CostExtra &extra, const Info<W> &info) {
    float startup_cost = 0.0f, run_cost = 0.0f, cpu_per_tuple, inner_run_cost, inner_rescan_run_cost;
    float outer_path_rows = outer.rows, inner_path_rows = inner.rows, ntuples;
    if (!info.params.enable_nestloop) startup_cost += info.params.disable_cost;
    startup_cost += inner.cost.startup + outer.cost.startup;
    run_cost += outer.cost.total - outer.cost.startup;
    if (outer_path_rows > 1.0f) run_cost += (outer_path_rows - 1.0f) * inner.cost.startup;
    inner_run_cost = inner.cost.total - inner.cost.startup;
    inner_rescan_run_cost = inner_run_cost;
    if (extra.inner_unique) {
        float outer_matched_rows = rintf(outer_path_rows * extra.outer_match_frac);
        float outer_unmatched_rows = outer_path_rows - outer_matched_rows;
        float inner_scan_frac = 2.0f / (extra.match_count + 1.0f);
        ntuples = outer_matched_rows * inner_path_rows * inner_scan_frac;
        if (extra.indexed_join_quals) {
            run_cost += inner_run_cost * inner_scan_frac;
            if (outer_matched_rows > 1.0f)
                run_cost += (outer_matched_rows - 1.0f) * inner_rescan_run_cost * inner_scan_frac;
            run_cost += outer_unmatched_rows * inner_rescan_run_cost / inner_path_rows;
        } else {
            ntuples += outer_unmatched_rows * inner_path_rows;
            run_cost += inner_run_cost;
            if (outer_unmatched_rows >= 1.0f) outer_unmatched_rows -= 1.0f;
            else outer_matched_rows -= 1.0f;
            if (outer_matched_rows > 0.0f)
                run_cost += outer_matched_rows * inner_rescan_run_cost * inner_scan_frac;
            if (outer_unmatched_rows > 0.0f)
                run_cost += outer_unmatched_rows * inner_rescan_run_cost;
        }
    } else {
        run_cost += inner_run_cost;
        if (outer_path_rows > 1.0f) run_cost += (outer_path_rows - 1.0f) * inner_rescan_run_cost;
        ntuples = outer_path_rows * inner_path_rows;
    }
    QualCost q = cost_qual_eval(inner_id, outer_id, info);
    startup_cost += q.startup;
    cpu_per_tuple = info.params.cpu_tuple_cost + q.per_tuple;
    run_cost += cpu_per_tuple * ntuples;
    PathCost pc; pc.startup = startup_cost; pc.total = startup_cost + run_cost;
    return pc;
}

/* cost_hashjoin: cost_postgres.cuh:287-470 */
template <int W>
static PathCost cost_hashjoin(const Bits<W> &outer_id, const JoinRel &outer, const Bits<W> &inner_id,
                              const JoinRel &inner, CostExtra extra, const Info<W> &info) {
    float startup_cost = 0.0f, run_cost = 0.0f, cpu_per_tuple, hashjointuples, virtualbuckets;
    float outer_path_rows = outer.rows, inner_path_rows = inner.rows;
    float inner_path_rows_total = inner_path_rows;
    if (!info.params.enable_hashjoin) startup_cost += info.params.disable_cost;
    QualCost hq = cost_qual_eval(outer_id, inner_id, info);
    startup_cost += outer.cost.startup;
    run_cost += outer.cost.total - outer.cost.startup;
    startup_cost += inner.cost.total;
    startup_cost += (info.params.cpu_operator_cost * (float)hq.n_quals + info.params.cpu_tuple_cost) * inner_path_rows;
    run_cost += info.params.cpu_operator_cost * (float)hq.n_quals * outer_path_rows;
    HtszRes h = exec_choose_hash_table_size(inner_path_rows_total, inner.width, info.params.work_mem);
    if (h.numbatches > 1) {
        float outerpages = page_size(outer_path_rows, outer.width);
        float innerpages = page_size(inner_path_rows, inner.width);
        startup_cost += info.params.seq_page_cost * innerpages;
        run_cost += info.params.seq_page_cost * (innerpages + 2.0f * outerpages);
    }
    virtualbuckets = (float)h.numbuckets * (float)h.numbatches;
    BucketStats bs = estimate_hash_bucket_stats(outer_id, inner_id, float_to_int_x86(virtualbuckets), info);
    if (relation_byte_size(clamp_row_est(inner_path_rows * bs.mcvfreq), inner.width) >
        (float)(info.params.work_mem * 1024L))
        startup_cost += info.params.disable_cost;
    if (extra.inner_unique) {
        float outer_matched_rows = rintf(outer_path_rows * extra.outer_match_frac);
        float inner_scan_frac = 2.0f / (extra.match_count + 1.0f);
        startup_cost += hq.startup;
        run_cost += hq.per_tuple * outer_matched_rows *
                    clamp_row_est(inner_path_rows * bs.bucketsize * inner_scan_frac) * 0.5f;
        run_cost += hq.per_tuple * (outer_path_rows - outer_matched_rows) *
                    clamp_row_est(inner_path_rows / virtualbuckets) * 0.05f;
        hashjointuples = outer_matched_rows;
    } else {
        startup_cost += hq.startup;
        run_cost += hq.per_tuple * outer_path_rows * clamp_row_est(inner_path_rows * bs.bucketsize) * 0.5f;
        hashjointuples = extra.joinrows;
    }
    cpu_per_tuple = info.params.cpu_tuple_cost;
    run_cost += cpu_per_tuple * hashjointuples;
    PathCost pc; pc.startup = startup_cost; pc.total = startup_cost + run_cost;
    return pc;
}

/* is_inner_unique: gpuqo_cost.cuh:43-74 */
template <int W>
static bool is_inner_unique(const Bits<W> &outer, const Bits<W> &inner, const Info<W> &info) {
    if (inner.size() != 1) return false;
    for (int i = 0, off_vars = 0; i < info.n_ec; off_vars += info.ec_relids[i].size(), i++) {
        const Bits<W> &ec = info.ec_relids[i];
        Bits<W> ml = ec & outer, mr = ec & inner;
        if (ml.empty() || mr.empty()) continue;
        int in_idx = member_idx(mr, ec);
        const gq_var_info &v = info.ec_vars[off_vars + in_idx];
        if (v.index_available && v.index_unique) return true;
    }
    return false;
}

/* The simple cost function (gpuqo_cost_simple.cuh:26-74): BASEREL_COEFF 0.2,
 * HASHJOIN_COEFF 1.0; every startup cost is 0. */
static const float BASEREL_COEFF_ = 0.2f, HASHJOIN_COEFF_ = 1.0f;

/* cost_baserel: cost_postgres.cuh:843-847 (the relation's own cost),
 * cost_simple.cuh:30-37, cost_cout.cuh:26-33 */
template <int W>
static PathCost cost_baserel(const BaseRel<W> &b, const Params &p) {
    if (p.cost_function == 1) return PathCost{0.0f, b.composite ? b.cost.total : BASEREL_COEFF_ * b.tuples};
    if (p.cost_function == 2) return PathCost{0.0f, b.composite ? b.cost.total : 0.0f};
    return b.cost;
}

/* cost_nestloop and cost_hashjoin of the build's cost function:
 * simple (cost_simple.cuh:44-56, :58-70), C_out (cost_cout.cuh:40-52,
 * :54-66), postgres (above) */
template <int W>
static PathCost nestloop_of(const Bits<W> &outer_id, const JoinRel &outer, const Bits<W> &inner_id,
                            const JoinRel &inner, const CostExtra &extra, const Info<W> &info) {
    if (info.params.cost_function == 1) return PathCost{0.0f, outer.cost.total + outer.rows * inner.cost.total};
    if (info.params.cost_function == 2) return PathCost{0.0f, INFF_};
    return cost_nestloop(outer_id, outer, inner_id, inner, extra, info);
}
template <int W>
static PathCost hashjoin_of(const Bits<W> &outer_id, const JoinRel &outer, const Bits<W> &inner_id,
                            const JoinRel &inner, const CostExtra &extra, const Info<W> &info) {
    if (info.params.cost_function == 1)
        return PathCost{0.0f, HASHJOIN_COEFF_ * extra.joinrows + inner.cost.total + outer.cost.total};
    if (info.params.cost_function == 2) return PathCost{0.0f, extra.joinrows + inner.cost.total + outer.cost.total};
    return cost_hashjoin(outer_id, outer, inner_id, inner, extra, info);
}

/* calc_join_cost: gpuqo_cost.cuh:76-158 */
template <int W>
static PathCost calc_join_cost(const Bits<W> &outer_id, const JoinRel &outer, const Bits<W> &inner_id,
                               const JoinRel &inner, float join_rows, const Info<W> &info) {
    PathCost min_cost, tmp;
    CostExtra extra;
    extra.inner_unique = is_inner_unique(outer_id, inner_id, info);
    extra.indexed_join_quals = false;
    extra.joinrows = join_rows;
    extra.outer_match_frac = join_rows / (outer.rows * inner.rows);
    extra.match_count = inner.rows;
    min_cost.startup = 0.0f;
    min_cost.total = NANF_;
    tmp = nestloop_of(outer_id, outer, inner_id, inner, extra, info);
    if (std::isnan(min_cost.total) || tmp.total < min_cost.total) min_cost = tmp;
    if (info.params.enable_hashjoin) {
        tmp = hashjoin_of(outer_id, outer, inner_id, inner, extra, info);
        if (tmp.total < min_cost.total) min_cost = tmp;
    }
    extra.indexed_join_quals = true;
    if (inner_id.size() == 1) {
        const BaseRel<W> &br = info.base_rels[inner_id.lowestPos() - 1];
        if (!br.composite) {
            for (int i = 0, off_vars = 0; i < info.n_ec; off_vars += info.ec_relids[i].size(), i++) {
                const Bits<W> &ec = info.ec_relids[i];
                Bits<W> ml = ec & outer_id, mr = ec & inner_id;
                if (ml.empty() || mr.empty()) continue;
                int in_idx = member_idx(mr, ec);
                const gq_var_info &v = info.ec_vars[off_vars + in_idx];
                if (!v.index_available) continue;
                JoinRel ir;
                ir.rows = v.index_rows;
                ir.cost.startup = v.index_cost.startup;
                ir.cost.total = v.index_cost.total;
                ir.width = inner.width;   /* not read by cost_nestloop */
                tmp = nestloop_of(outer_id, outer, inner_id, ir, extra, info);
                if (tmp.total < min_cost.total) min_cost = tmp;
            }
        }
    }
    return min_cost;
}

/* fkselc_of_ec: row_estimation.cuh:26-45 */
template <int W>
static float fkselc_of_ec(int off_fk, const Bits<W> &ec, const Bits<W> &outer_rels, const Bits<W> &inner_rels,
                          const Info<W> &info) {
    Bits<W> tmp = outer_rels;
    while (!tmp.empty()) {
        Bits<W> out_id = tmp.lowest();
        int out_idx = (out_id.allLower() & ec).size();
        Bits<W> match = inner_rels & info.ec_fks[off_fk + out_idx];
        if (!match.empty()) {
            int in_idx = match.lowestPos() - 1;
            return (float)(1.0 / (double)fmaxf(1.0f, info.base_rels[in_idx].tuples));
        }
        tmp -= out_id;
    }
    return NANF_;
}

/* estimate_ec_selectivity: row_estimation.cuh:48-86 */
template <int W>
static float estimate_ec_selectivity(const Bits<W> &ec, int off_sels, int off_fks, const Bits<W> &l,
                                     const Bits<W> &r, const Info<W> &info) {
    int size = ec.size();
    Bits<W> ml = ec & l, mr = ec & r;
    if (ml.empty() || mr.empty()) return 1.0f;
    float fk = fkselc_of_ec(off_fks, ec, ml, mr, info);
    if (!std::isnan(fk)) return fk;
    fk = fkselc_of_ec(off_fks, ec, mr, ml, info);
    if (!std::isnan(fk)) return fk;
    int idx_l = (ml.allLower() & ec).size();
    int idx_r = (mr.allLower() & ec).size();
    return info.ec_sels[off_sels + eqClassIndex(idx_l, idx_r, size)];
}

/* estimate_join_selectivity: row_estimation.cuh:90-117 */
template <int W>
static float estimate_join_selectivity(const Bits<W> &l, const Bits<W> &r, const Info<W> &info) {
    float sel = 1.0f;
    int off_sels = 0, off_fks = 0;
    for (int i = 0; i < info.n_ec; i++) {
        const Bits<W> &ec = info.ec_relids[i];
        sel *= estimate_ec_selectivity(ec, off_sels, off_fks, l, r, info);
        int s = ec.size();
        off_sels += eqClassNSels(s);
        off_fks += s;
    }
    return sel;
}

/* estimate_join_rows: row_estimation.cuh:120-147 */
template <int W>
static float estimate_join_rows(const Bits<W> &l, float lrows, const Bits<W> &r, float rrows, const Info<W> &info) {
    float sel = estimate_join_selectivity(l, r, info);
    float rows = sel * lrows * rrows;
    return rows > 1 ? roundf(rows) : 1.0f;
}

/* make_join_relation: gpuqo_cpu_common.cuh:168-186 */
template <int W>
static JoinRel make_join(const Bits<W> &l, const JoinRel &L, const Bits<W> &r, const JoinRel &R,
                         const Info<W> &info) {
    JoinRel j;
    j.rows = estimate_join_rows(l, L.rows, r, R.rows, info);
    j.cost = calc_join_cost(l, L, r, R, j.rows, info);
    j.width = L.width + R.width;   /* get_join_width, row_estimation.cuh:150-158 */
    return j;
}

/* ------------------------------------------------------------------------
 * Connectivity: gpuqo_filter.cuh:113-163
 * ---------------------------------------------------------------------- */
template <int W>
static Bits<W> grow(Bits<W> from, const Bits<W> &subset, const std::vector<Bits<W>> &et) {
    Bits<W> V, N = from;
    while (!N.empty()) {
        int idx = N.lowestPos() - 1;
        V.set(idx + 1);
        N |= et[idx];
        N &= subset;
        N -= V;
    }
    return V;
}
template <int W>
static bool is_connected(const Bits<W> &s, const std::vector<Bits<W>> &et) {
    if (s.empty()) return false;
    return grow(s.lowest(), s, et) == s;
}
/* get_neighbours: gpuqo_filter.cuh:93-105 */
template <int W>
static Bits<W> get_neighbours(const Bits<W> &set, const std::vector<Bits<W>> &et) {
    Bits<W> n, t = set;
    while (!t.empty()) { int i = t.lowestPos() - 1; n |= et[i]; t.unset(i + 1); }
    return n - set;
}

/* ------------------------------------------------------------------------
 * Remapper: gpuqo_remapper.cuh/.cu
 * ---------------------------------------------------------------------- */
template <int W>
struct RemapEl {
    Bits<W> from_relid;
    int to_idx;
    QueryTree<W> *qt;
};

template <int WI, int WO>
struct Remapper {
    std::list<RemapEl<WI>> transf;
    explicit Remapper(const std::list<RemapEl<WI>> &t) : transf(t) {}

    Bits<WO> remapRelid(const Bits<WI> &id) const {                 /* remapper.cu:43-54 */
        Bits<WO> out;
        for (auto &e : transf) if (e.from_relid.intersects(id)) out.set(e.to_idx + 1);
        return out;
    }
    Bits<WO> remapRelidNoComposite(const Bits<WI> &id, const Info<WI> &from) const {   /* :56-70 */
        Bits<WO> out;
        for (auto &e : transf) {
            if (e.from_relid.size() > 1 || from.base_rels[e.from_relid.lowestPos() - 1].composite) continue;
            if (e.from_relid.intersects(id)) out.set(e.to_idx + 1);
        }
        return out;
    }
    Bits<WI> remapRelidInv(const Bits<WO> &id) const {              /* :72-82 */
        Bits<WI> out;
        for (auto &e : transf) if (id.isSet(e.to_idx + 1)) out |= e.from_relid;
        return out;
    }
    bool ecDropped(const Bits<WI> &ec) const {                      /* countEqClasses :17-41 */
        for (auto &e : transf) if (ec.isSubset(e.from_relid)) return true;
        return false;
    }

    Info<WO> *remapPlannerInfo(const Info<WI> &old) const {         /* :200-276 */
        Info<WO> *info = new Info<WO>();
        int n = (int)transf.size();
        info->n_rels = n;
        info->n_iters = old.n_iters;
        info->params = old.params;
        info->base_rels.resize(n);
        info->edge_table.resize(n);
        /* remapEdgeTable :84-107 */
        for (auto &e : transf) {
            Bits<WO> et;
            Bits<WI> t = e.from_relid;
            while (!t.empty()) {
                int fi = t.lowestPos() - 1;
                et |= remapRelid(old.edge_table[fi]);
                t.unset(fi + 1);
            }
            et.unset(e.to_idx + 1);
            info->edge_table[e.to_idx] = et;
        }
        /* remapBaseRels :109-140 */
        for (auto &e : transf) {
            BaseRel<WO> &b = info->base_rels[e.to_idx];
            b.id = remapRelid(e.from_relid);
            if (e.qt) {
                b.rows = e.qt->rows; b.cost = e.qt->cost; b.width = e.qt->width; b.tuples = e.qt->rows;
            } else {
                const BaseRel<WI> &o = old.base_rels[e.from_relid.lowestPos() - 1];
                b.rows = o.rows; b.cost = o.cost; b.width = o.width; b.tuples = o.tuples;
            }
            if (e.from_relid.size() == 1) b.composite = old.base_rels[e.from_relid.lowestPos() - 1].composite;
            else b.composite = true;
        }
        /* equivalence classes :237-273, remapEqClass :142-198 */
        int old_off_sel = 0, old_off_fk = 0;
        for (int i = 0; i < old.n_ec; i++) {
            const Bits<WI> &ecf = old.ec_relids[i];
            if (!ecDropped(ecf)) {
                Bits<WO> ect = remapRelid(ecf);
                int s_from = ecf.size(), s_to = ect.size();
                size_t sel_base = info->ec_sels.size();
                size_t fk_base = info->ec_fks.size();
                info->ec_sels.resize(sel_base + eqClassNSels(s_to), 0.0f);
                info->ec_fks.resize(fk_base + s_to, Bits<WO>());
                info->ec_vars.resize(fk_base + s_to);
                for (int lt = 0; lt < s_to; lt++) {
                    Bits<WO> id_l_to = expandToMask(lt, ect);
                    Bits<WI> id_l_from = remapRelidInv(id_l_to);
                    int lf = (id_l_from.allLower() & ecf).size();
                    for (int rt = lt + 1; rt < s_to; rt++) {
                        Bits<WO> id_r_to = expandToMask(rt, ect);
                        Bits<WI> id_r_from = remapRelidInv(id_r_to);
                        int rf = (id_r_from.allLower() & ecf).size();
                        int to_i = eqClassIndex(lt, rt, s_to);
                        int from_i = eqClassIndex(lf, rf, s_from);
                        if (id_l_from.size() == id_l_to.size() && id_r_from.size() == id_r_to.size())
                            info->ec_sels[sel_base + to_i] = old.ec_sels[old_off_sel + from_i];
                        else
                            info->ec_sels[sel_base + to_i] =
                                estimate_ec_selectivity(ecf, old_off_sel, old_off_fk, id_l_from, id_r_from, old);
                    }
                    if (id_l_from.size() == 1 && !old.base_rels[id_l_from.lowestPos() - 1].composite)
                        info->ec_fks[fk_base + lt] = remapRelidNoComposite(old.ec_fks[old_off_fk + lf], old);
                    info->ec_vars[fk_base + lt] = old.ec_vars[old_off_fk + lf];
                    if (id_l_from.size() > 1) info->ec_vars[fk_base + lt].index_available = 0;
                }
                info->ec_relids.push_back(ect);
                info->n_ec++;
            }
            old_off_fk += ecf.size();
            old_off_sel += eqClassNSels(ecf.size());
        }
        return info;
    }

    QueryTree<WI> *remapQueryTree(QueryTree<WO> *qt) const {        /* :278-301 */
        if (!qt) return nullptr;
        if (qt->id.size() == 1) {
            int idx = qt->id.lowestPos() - 1;
            for (auto &e : transf) if (e.qt && e.to_idx == idx) return e.qt;
        }
        QueryTree<WI> *o = new QueryTree<WI>();
        o->id = remapRelidInv(qt->id);
        o->left = remapQueryTree(qt->left);
        o->right = remapQueryTree(qt->right);
        o->rows = qt->rows; o->cost = qt->cost; o->width = qt->width;
        return o;
    }
    QueryTree<WO> *remapQueryTreeFwd(QueryTree<WI> *qt) const {     /* :303-321 */
        if (!qt) return nullptr;
        QueryTree<WO> *o = new QueryTree<WO>();
        o->id = remapRelid(qt->id);
        if (o->id.size() == 1) { o->left = nullptr; o->right = nullptr; }
        else { o->left = remapQueryTreeFwd(qt->left); o->right = remapQueryTreeFwd(qt->right); }
        o->rows = qt->rows; o->cost = qt->cost; o->width = qt->width;
        return o;
    }
};

/* makeBFSIndexRemapper: gpuqo_bfs_indexing.cu:13-55 */
template <int W>
static std::list<RemapEl<W>> bfs_remap_list(const Info<W> &info) {
    std::vector<int> q(info.n_rels + 1);
    int front = 0, back = 0, bfs_idx = 0;
    std::list<RemapEl<W>> lst;
    q[back++] = 0;
    Bits<W> seen = Bits<W>::nth(1);
    while (front != back && bfs_idx < info.n_rels) {
        int u = q[front++];
        Bits<W> edges = info.edge_table[u];
        RemapEl<W> el; el.from_relid = Bits<W>::nth(u + 1); el.to_idx = bfs_idx; el.qt = nullptr;
        lst.push_back(el);
        bfs_idx++;
        while (!edges.empty()) {
            int next = edges.lowestPos();
            if (!seen.isSet(next)) q[back++] = next - 1;
            edges.unset(next);
        }
        seen |= info.edge_table[u];
    }
    return lst;
}

template <int W>
static std::list<RemapEl<W>> identity_list(const Info<W> &info) {
    std::list<RemapEl<W>> lst;
    for (int i = 0; i < info.n_rels; i++) {
        RemapEl<W> el; el.from_relid = info.base_rels[i].id; el.to_idx = i; el.qt = nullptr;
        lst.push_back(el);
    }
    return lst;
}

/* ------------------------------------------------------------------------
 * Sequential CPU DP framework: gpuqo_cpu_common.cuh, gpuqo_cpu_sequential.cu
 * ---------------------------------------------------------------------- */
template <int W>
struct JR {                         /* JoinRelationCPU */
    Bits<W> id, left_id, right_id, edges;
    JoinRel v;
    JR *lp, *rp;
};

template <int W>
using Memo = std::unordered_map<Bits<W>, JR<W> *, BitsHash<W>>;

struct Counters { uint64_t checks = 0, joins = 0, sets = 0; };

template <int W>
static void build_query_tree(const JR<W> *jr, QueryTree<W> **qt) {  /* cpu_common.cuh:114-133 */
    if (!jr) { *qt = nullptr; return; }
    QueryTree<W> *t = new QueryTree<W>();
    t->id = jr->id; t->rows = jr->v.rows; t->cost = jr->v.cost; t->width = jr->v.width;
    t->left = t->right = nullptr;
    *qt = t;
    build_query_tree(jr->lp, &t->left);
    build_query_tree(jr->rp, &t->right);
}

/* do_join, keeping the first entry unless strictly cheaper:
 * gpuqo_cpu_common.cuh:189-211.  Returns true when the set is new. */
template <int W>
static bool do_join(JR<W> &L, JR<W> &R, const Info<W> &info, Memo<W> &memo, JR<W> **out, Counters &c) {
    JR<W> j;
    j.id = L.id | R.id;
    j.left_id = L.id; j.right_id = R.id; j.lp = &L; j.rp = &R;
    j.edges = L.edges | R.edges;
    j.v = make_join(L.id, L.v, R.id, R.v, info);
    c.joins++;
    auto it = memo.find(j.id);
    if (it != memo.end()) {
        if (j.v.cost.total < it->second->v.cost.total) *it->second = j;
        *out = it->second;
        return false;
    }
    JR<W> *n = new JR<W>(j);
    memo.insert(std::make_pair(j.id, n));
    *out = n;
    return true;
}

template <int W>
static void init_memo(const Info<W> &info, Memo<W> &memo) {     /* cpu_sequential.cu:73-85 */
    for (int i = 0; i < info.n_rels; i++) {
        JR<W> *jr = new JR<W>();
        jr->id = info.base_rels[i].id;
        jr->lp = jr->rp = nullptr;
        jr->v.cost = cost_baserel(info.base_rels[i], info.params);
        jr->v.width = info.base_rels[i].width;
        jr->v.rows = info.base_rels[i].rows;
        jr->edges = info.edge_table[i];
        memo.insert(std::make_pair(info.base_rels[i].id, jr));
    }
}

/* final set selection + tree: cpu_sequential.cu:96-118 */
template <int W>
static QueryTree<W> *finish_memo(const Info<W> &info, Memo<W> &memo) {
    Bits<W> final_id;
    if (info.n_rels == info.n_iters) {
        for (int i = 0; i < info.n_rels; i++) final_id |= info.base_rels[i].id;
    } else {
        float min_cost = INFF_;
        for (auto it = memo.begin(); it != memo.end(); ++it)
            if (it->first.size() == info.n_iters && it->second->v.cost.total < min_cost) {
                min_cost = it->second->v.cost.total;
                final_id = it->first;
            }
    }
    QueryTree<W> *out = nullptr;
    auto f = memo.find(final_id);
    if (f != memo.end()) build_query_tree(f->second, &out);
    for (auto it = memo.begin(); it != memo.end(); ++it) delete it->second;
    return out;
}

/* DPsubGenericCPUAlgorithm::enumerate: gpuqo_cpu_dpsub.cuh:26-63.
 * visit(S) is called for each connected S in the reference's order. */
template <int W, typename F>
static void dpsub_enumerate_sets(const Info<W> &info, F visit) {
    if (info.n_iters == info.n_rels) {
        uint64_t lim = 1ULL << info.n_rels;
        for (uint64_t i = 1; i < lim; i++) {
            Bits<W> s(i << 1);
            if (!is_connected(s, info.edge_table)) continue;
            visit(s);
        }
    } else {
        for (int i = 2; i <= info.n_iters; i++) {
            uint64_t from = 0, to = 0;
            for (int j = 0; j < i; j++) { from |= 1ULL << j; to |= 1ULL << (info.n_rels - 1 - j); }
            uint64_t s = from;
            while (true) {
                Bits<W> js(s << 1);
                if (is_connected(js, info.edge_table)) visit(js);
                if (s != to) s = next_perm(s); else break;
            }
        }
    }
}

/* The connected sets of sizes 2 .. n_iters, size by size, each size in
 * increasing order: every neighbour v outside a set S of size k gives
 * S + v of size k + 1 (duplicates dropped).  A checker-side stand-in for
 * dpsub_enumerate_sets on DPs whose pass over all 2^n integers would not
 * finish (the wide DPs above 32 relations).  It visits the same connected
 * sets, each once, and every set after all of its proper subsets (sizes
 * ascend).  A set's entry depends only on its splits, and their entries are
 * final when it is visited, so the memo ends identical
 * (oracle_plan_connected; pinned against the 2^n pass on small graphs). */
template <int W, typename F>
static void connected_enumerate_sets(const Info<W> &info, F visit) {
    std::vector<Bits<W>> cur, nxt;
    for (int i = 0; i < info.n_rels; i++) cur.push_back(Bits<W>::nth(i + 1));
    for (int k = 1; k < info.n_iters; k++) {
        nxt.clear();
        for (const Bits<W> &S : cur) {
            Bits<W> nb;
            for (Bits<W> t = S; !t.empty();) {
                const int u = t.lowestPos();
                t.unset(u);
                nb |= info.edge_table[u - 1];
            }
            nb -= S;
            while (!nb.empty()) {
                const int v = nb.lowestPos();
                nb.unset(v);
                Bits<W> T = S;
                T.set(v);
                nxt.push_back(T);
            }
        }
        std::sort(nxt.begin(), nxt.end());
        nxt.erase(std::unique(nxt.begin(), nxt.end()), nxt.end());
        cur.swap(nxt);
        for (const Bits<W> &S : cur) visit(S);
    }
}

template <int W>
static QueryTree<W> *cpu_dpsub_bicc(const Info<W> &info, Counters &c, bool connected_only = false);
template <int W, typename F>
static void bicc_pairs(const Bits<W> &S, const std::vector<Bits<W>> &et, Counters &c, F emit);

/* cpu_dpsub over connected_enumerate_sets: the valid splits of a set are
 * those with both sides connected, which are exactly the block splits
 * (bicc_pairs: two cut edges in different blocks would close a cycle
 * through both), taken in cpu_dpsub's order, left side ascending.  c.checks
 * still counts the 2^|S| - 2 subsets cpu_dpsub tries. */
template <int W>
static QueryTree<W> *cpu_dpsub_connected(const Info<W> &info, Counters &c) {
    Memo<W> memo;
    init_memo(info, memo);
    std::vector<std::pair<Bits<W>, Bits<W>>> ps;
    connected_enumerate_sets(info, [&](const Bits<W> &S) {
        c.sets++;
        c.checks += (1ULL << S.size()) - 2ULL;
        ps.clear();
        Counters unused;
        bicc_pairs(S, info.edge_table, unused, [&](const Bits<W> &l, const Bits<W> &r) { ps.emplace_back(l, r); });
        std::sort(ps.begin(), ps.end(),
                  [](const std::pair<Bits<W>, Bits<W>> &a, const std::pair<Bits<W>, Bits<W>> &b) {
                      return a.first < b.first;
                  });
        for (const auto &p : ps) {
            auto lf = memo.find(p.first), rf = memo.find(p.second);
            JR<W> *o;
            do_join(*lf->second, *rf->second, info, memo, &o, c);
        }
    });
    return finish_memo(info, memo);
}

/* cpu_dpsub: DPsubCPUAlgorithm::enumerate_subsets, gpuqo_cpu_dpsub.cu:32-79 */
template <int W>
static QueryTree<W> *cpu_dpsub(const Info<W> &info, Counters &c, bool connected_only = false) {
    static_assert(W == 1, "DPsub oracle works on <64 relations");
    if (connected_only) return cpu_dpsub_connected(info, c);
    Memo<W> memo;
    init_memo(info, memo);
    dpsub_enumerate_sets(info, [&](const Bits<W> &S) {
        if (S.size() > 1) c.sets++;
        Bits<W> l = S.lowest();
        while (l != S) {
            Bits<W> r = S - l;
            if (!l.empty() && !r.empty()) {
                c.checks++;
                auto lf = memo.find(l), rf = memo.find(r);
                if (lf != memo.end() && rf != memo.end()) {
                    JR<W> *o;
                    do_join(*lf->second, *rf->second, info, memo, &o, c);
                }
            }
            l = nextSubset(l, S);
        }
    });
    return finish_memo(info, memo);
}

/* DPsubBiCCCPUAlgorithm::find_blocks: gpuqo_cpu_dpsub_bicc.cu:65-169.
 * Iterative DFS from the lowest vertex, neighbours in ascending order; a
 * block is emitted when a tree edge (u,v) finishes with low[v] >= depth[u]. */
template <int W>
static void find_blocks(const Bits<W> &set, const std::vector<Bits<W>> &et, std::vector<Bits<W>> &blocks) {
    struct Frame { int u; Bits<W> edges; bool remain; };
    struct Edge { int u, v; };
    std::vector<Frame> cs;
    std::vector<Edge> es;
    std::vector<char> visited(Bits<W>::SIZE, 0);
    std::vector<int> depth(Bits<W>::SIZE, -1), low(Bits<W>::SIZE, INT_MAX);
    std::vector<int> parent(Bits<W>::SIZE, 0);
    cs.push_back(Frame{set.lowestPos(), Bits<W>(), false});
    int curr_depth = 1;
    while (!cs.empty()) {
        Frame el = cs.back(); cs.pop_back();
        if (el.edges.empty()) {
            visited[el.u] = 1;
            depth[el.u] = curr_depth++;
            Bits<W> edges = et[el.u - 1] & set;
            if (!edges.empty()) cs.push_back(Frame{el.u, edges, false});
        } else {
            int v = el.edges.lowestPos();
            if (el.remain) {
                if (low[v] >= depth[el.u]) {
                    Bits<W> block;
                    Edge e;
                    do { e = es.back(); es.pop_back(); block.set(e.u); block.set(e.v); }
                    while (e.u != el.u || e.v != v);
                    blocks.push_back(block);
                }
                low[el.u] = std::min(low[el.u], low[v]);
                el.edges.unset(v);
                if (!el.edges.empty()) cs.push_back(Frame{el.u, el.edges, false});
            } else if (!visited[v]) {
                es.push_back(Edge{el.u, v});
                parent[v] = el.u;
                cs.push_back(Frame{el.u, el.edges, true});
                cs.push_back(Frame{v, Bits<W>(), false});
            } else if (parent[el.u] != v && depth[v] < depth[el.u]) {
                es.push_back(Edge{el.u, v});
                low[el.u] = std::min(low[el.u], depth[v]);
                el.edges.unset(v);
                if (!el.edges.empty()) cs.push_back(Frame{el.u, el.edges, false});
            } else {
                el.edges.unset(v);
                if (!el.edges.empty()) cs.push_back(Frame{el.u, el.edges, false});
            }
        }
    }
}

/* enumerate_subsets of the BiCC variant: gpuqo_cpu_dpsub_bicc.cu:172-236.
 * Calls emit(left, right) for every valid pair of S in the reference's order. */
template <int W, typename F>
static void bicc_pairs(const Bits<W> &S, const std::vector<Bits<W>> &et, Counters &c, F emit) {
    std::vector<Bits<W>> blocks;
    find_blocks(S, et, blocks);
    for (const Bits<W> &block : blocks) {
        Bits<W> lb = block.lowest();
        while (lb != block) {
            Bits<W> rb = block - lb;
            Bits<W> left = grow(lb.lowest(), S - rb, et);
            Bits<W> right = grow(rb.lowest(), S - lb, et);
            c.checks++;
            if ((left | right) == S) emit(left, right);
            lb = nextSubset(lb, block);
        }
    }
}

template <int W>
static QueryTree<W> *cpu_dpsub_bicc(const Info<W> &info, Counters &c, bool connected_only) {
    static_assert(W == 1, "DPsub oracle works on <64 relations");
    Memo<W> memo;
    init_memo(info, memo);
    auto per_set = [&](const Bits<W> &S) {
        if (S.size() > 1) c.sets++;
        bicc_pairs(S, info.edge_table, c, [&](const Bits<W> &l, const Bits<W> &r) {
            auto lf = memo.find(l), rf = memo.find(r);
            JR<W> *o;
            do_join(*lf->second, *rf->second, info, memo, &o, c);
        });
    };
    if (connected_only) connected_enumerate_sets(info, per_set);
    else dpsub_enumerate_sets(info, per_set);
    return finish_memo(info, memo);
}

/* cpu_dpsize: DPsizeCPUAlgorithm, gpuqo_cpu_dpsize.cu:27-93 */
template <int W>
static QueryTree<W> *cpu_dpsize(const Info<W> &info, Counters &c) {
    Memo<W> memo;
    init_memo(info, memo);
    std::vector<std::vector<JR<W> *>> levels(info.n_rels + 1);
    for (auto it = memo.begin(); it != memo.end(); ++it) levels[1].push_back(it->second);
    for (int js = 2; js <= info.n_iters; js++) {
        for (int ls = js - 1; ls >= 1; ls--) {
            int rs = js - ls;
            /* indices, not iterators: level js grows while level ls/rs < js is read */
            for (size_t a = 0; a < levels[ls].size(); a++) {
                for (size_t b = 0; b < levels[rs].size(); b++) {
                    JR<W> &L = *levels[ls][a], &R = *levels[rs][b];
                    c.checks++;
                    if (L.id.intersects(R.id) || !L.edges.intersects(R.id)) continue;  /* check_join :75-81 */
                    JR<W> *o;
                    if (do_join(L, R, info, memo, &o, c)) levels[js].push_back(o);   /* post_join :83-92 */
                }
            }
        }
    }
    for (int k = 2; k <= info.n_rels; k++) c.sets += levels[k].size();
    return finish_memo(info, memo);
}

/* ------------------------------------------------------------------------
 * parallel_cpu_dpsub[_bicc]: gpuqo_cpu_dpsub_parallel.cu:48-292 and
 * gpuqo_cpu_level_hashtable.cuh.  Timed baseline; the per-set order
 * (manage_best) equals the sequential one, so the plans are identical.
 * ---------------------------------------------------------------------- */
template <int W>
struct LevelMemo {
    std::vector<Memo<W>> buckets;
    explicit LevelMemo(int n) : buckets(n + 1) {}
    JR<W> *find(const Bits<W> &k) {
        auto &b = buckets[k.size()];
        auto it = b.find(k);
        return it == b.end() ? nullptr : it->second;
    }
};

static uint64_t binom_u64(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return r;
}

/* dpsub_unrank_sid: gpuqo_dpsub.cuh:95-114 (rank -> k-subset, numeric order) */
static uint64_t unrank_sid(uint64_t sid, int qss, int sq) {
    uint64_t s = (sq >= 64) ? ~0ULL : ((1ULL << sq) - 1);
    int q = qss, t = sq;
    while (t > 0 && t > q) {
        uint64_t o = binom_u64(t - 1, t - q - 1);
        if (sid < o) s &= ~(1ULL << (t - 1));
        else { q--; sid -= o; }
        t--;
    }
    return s;
}

template <int W>
static QueryTree<W> *parallel_cpu_dpsub(const Info<W> &info, bool bicc, int n_threads, int chunk, Counters &c) {
    static_assert(W == 1, "DPsub oracle works on <64 relations");
    LevelMemo<W> memo(info.n_rels);
    for (int i = 0; i < info.n_rels; i++) {
        JR<W> *jr = new JR<W>();
        jr->id = info.base_rels[i].id; jr->lp = jr->rp = nullptr;
        jr->v.cost = cost_baserel(info.base_rels[i], info.params); jr->v.width = info.base_rels[i].width;
        jr->v.rows = info.base_rels[i].rows; jr->edges = info.edge_table[i];
        memo.buckets[1].insert(std::make_pair(jr->id, jr));
    }
    /* Persistent workers with a per-level barrier (mutex + two condition
     * variables) and an atomic work counter handing out `skip` consecutive
     * ranks: gpuqo_cpu_dpsub_parallel.cu:48-107 and :145-222. */
    std::mutex mu;
    std::condition_variable start_cv, done_cv;
    int cur_level = 0, n_waiting = 0;
    uint64_t skip = 1;
    std::atomic<uint64_t> next_sid(0), n_sets_total(0), n_joins_total(0);
    auto worker = [&]() {
        int level = 0;
        while (level < info.n_iters) {
            {
                std::unique_lock<std::mutex> lk(mu);
                start_cv.wait(lk, [&] { return cur_level > level; });
                level = cur_level;
            }
            std::vector<std::pair<Bits<W>, JR<W> *>> mine;
            uint64_t n_sets = binom_u64(info.n_rels, level);
            uint64_t sid, local_sets = 0, local_joins = 0;
            while ((sid = next_sid.fetch_add(skip)) < n_sets) {
                uint64_t s = unrank_sid(sid, level, info.n_rels);
                for (uint64_t i = 0; sid + i < n_sets && i < skip; i++) {
                    Bits<W> set(s << 1);
                    if (is_connected(set, info.edge_table)) {
                        local_sets++;
                        JR<W> *best = nullptr;
                        auto consider = [&](const Bits<W> &l, const Bits<W> &r) {
                            JR<W> *L = memo.find(l), *R = memo.find(r);
                            if (!L || !R) return;
                            JR<W> j;
                            j.id = set; j.left_id = l; j.right_id = r; j.lp = L; j.rp = R;
                            j.edges = L->edges | R->edges;
                            j.v = make_join(l, L->v, r, R->v, info);
                            local_joins++;
                            if (!best) best = new JR<W>(j);
                            else if (j.v.cost.total < best->v.cost.total) *best = j;
                        };
                        if (bicc) {
                            Counters dummy;
                            bicc_pairs(set, info.edge_table, dummy, consider);
                        } else {
                            Bits<W> l = set.lowest();
                            while (l != set) {
                                Bits<W> r = set - l;
                                if (!l.empty() && !r.empty()) consider(l, r);
                                l = nextSubset(l, set);
                            }
                        }
                        mine.push_back(std::make_pair(set, best));
                    }
                    s = next_perm(s);
                }
            }
            std::lock_guard<std::mutex> g(mu);
            for (auto &p : mine) memo.buckets[level].insert(p);
            n_sets_total += local_sets;
            n_joins_total += local_joins;
            if (++n_waiting == n_threads) done_cv.notify_one();
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; t++) th.emplace_back(worker);
    {
        std::unique_lock<std::mutex> lk(mu);
        for (int level = 2; level <= info.n_iters; level++) {
            next_sid = 0;
            uint64_t n_sets = binom_u64(info.n_rels, level);
            skip = std::min((n_sets + n_threads - 1) / (uint64_t)n_threads, (uint64_t)chunk);
            cur_level = level;
            start_cv.notify_all();
            done_cv.wait(lk, [&] { return n_waiting == n_threads; });
            n_waiting = 0;
        }
        if (info.n_iters < 2) { cur_level = info.n_iters + 1; start_cv.notify_all(); }
    }
    for (auto &t : th) t.join();
    c.sets = n_sets_total; c.joins = n_joins_total;
    Bits<W> final_id;
    for (int i = 0; i < info.n_rels; i++) final_id |= info.base_rels[i].id;
    QueryTree<W> *out = nullptr;
    JR<W> *f = memo.find(final_id);
    if (f) build_query_tree(f, &out);
    for (auto &b : memo.buckets) for (auto &p : b) delete p.second;
    return out;
}

/* ------------------------------------------------------------------------
 * GOO: gpuqo_cpu_goo.cu:39-148
 * ---------------------------------------------------------------------- */
template <int W>
static bool are_valid_pair(const Bits<W> &l, const Bits<W> &r, const Info<W> &info) {  /* filter.cuh:60-90 */
    Bits<W> le, t = l;
    while (!t.empty()) { int i = t.lowestPos() - 1; le |= info.edge_table[i]; t.unset(i + 1); }
    return le.intersects(r) && !l.intersects(r);
}

template <int W>
static QueryTree<W> *cpu_goo(const Info<W> &info) {
    struct Cmp {
        bool operator()(QueryTree<W> *l, QueryTree<W> *r) const {
            if (l && r) return l->rows > r->rows;
            return r != nullptr;
        }
    };
    auto join = [&](QueryTree<W> &L, QueryTree<W> &R) {
        QueryTree<W> *qt = new QueryTree<W>();
        qt->id = L.id | R.id;
        qt->rows = estimate_join_rows(L.id, L.rows, R.id, R.rows, info);
        qt->cost.startup = 0;
        qt->cost.total = qt->rows + L.cost.total + R.cost.total;
        qt->left = &L; qt->right = &R;
        qt->width = L.width + R.width;
        return qt;
    };
    std::vector<QueryTree<W> *> heap;
    std::map<Bits<W>, QueryTree<W> *> rels;
    Cmp cmp;
    for (int i = 0; i < info.n_rels; i++) {
        QueryTree<W> *qt = new QueryTree<W>();
        qt->id = info.base_rels[i].id; qt->rows = info.base_rels[i].rows;
        qt->cost.startup = 0; qt->cost.total = info.base_rels[i].rows;
        qt->left = qt->right = nullptr; qt->width = info.base_rels[i].width;
        rels.insert(std::make_pair(qt->id, qt));
    }
    for (auto i = rels.begin(); i != rels.end(); i++)
        for (auto j = i; j != rels.end(); j++)
            if (are_valid_pair(i->first, j->first, info)) {
                heap.push_back(join(*i->second, *j->second));
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
    QueryTree<W> *smallest = nullptr;
    do {
        if (heap.empty()) return nullptr;   /* disconnected graph */
        std::pop_heap(heap.begin(), heap.end(), cmp);
        smallest = heap.back(); heap.pop_back();
        auto lp = rels.find(smallest->left->id), rp = rels.find(smallest->right->id);
        if (lp == rels.end() || rp == rels.end()) { delete smallest; smallest = nullptr; continue; }
        rels.erase(lp); rels.erase(rp);
        for (auto j = rels.begin(); j != rels.end(); j++)
            if (are_valid_pair(smallest->id, j->first, info)) {
                heap.push_back(join(*smallest, *j->second));
                std::push_heap(heap.begin(), heap.end(), cmp);
            }
        rels.insert(std::make_pair(smallest->id, smallest));
    } while (smallest == nullptr || smallest->id.size() < info.n_rels);
    for (QueryTree<W> *qt : heap) delete qt;
    /* Nodes referenced by `smallest` stay alive; the rest of `rels` only
     * holds `smallest` at this point. */
    return smallest;
}

/* ------------------------------------------------------------------------
 * Dispatch and drivers: gpuqo_main_internal.cu, gpuqo_idp.cu,
 * gpuqo_dpdp_union.cu
 * ---------------------------------------------------------------------- */
/* ------------------------------------------------------------------------
 * DPccp: gpuqo_cpu_dpccp.cu:27-103.  Connected subgraphs are grown from
 * each relation, highest first; for each one its complements are grown from
 * its neighbours, highest first.  `join(L, R)` receives every csg-cmp pair
 * once, in the reference's order, as emit() does (64-bit sets).
 * ---------------------------------------------------------------------- */
template <typename J>
struct DPccpEnum {
    const Info<1> &info;
    J &join;
    static Bits<1> lower_inc(const Bits<1> &s) {          /* bits | (bits - 1) */
        Bits<1> b;
        b.w[0] = s.w[0] | (s.w[0] - 1);
        return b;
    }
    static Bits<1> highest(const Bits<1> &s) {
        Bits<1> b;
        b.w[0] = s.w[0] ? (1ull << (63 - __builtin_clzll(s.w[0]))) : 0ull;
        return b;
    }
    void csg_rec(const Bits<1> &S, const Bits<1> &X, const Bits<1> &cmp) {   /* :31-49 */
        const Bits<1> N = get_neighbours(S, info.edge_table) - X;
        if (N.empty()) return;
        for (Bits<1> sub = N.lowest(); !sub.empty(); sub = nextSubset(sub, N)) emit(cmp, S | sub);
        for (Bits<1> sub = N.lowest(); !sub.empty(); sub = nextSubset(sub, N)) csg_rec(S | sub, X | N, cmp);
    }
    void cmp(const Bits<1> &S) {                           /* :60-78 */
        const Bits<1> X = lower_inc(S);
        const Bits<1> N = get_neighbours(S, info.edge_table) - X;
        Bits<1> t = N;
        while (!t.empty()) {
            const Bits<1> v = highest(t);
            emit(S, v);
            csg_rec(v, X | (lower_inc(v) & N), S);
            t -= v;
        }
    }
    void emit(const Bits<1> &l, const Bits<1> &r) {        /* :80-103 */
        if (!l.empty() && !r.empty()) join(l, r);
        else if (!l.empty()) cmp(l);
        else cmp(r);
    }
    void run() {                                           /* :51-58 */
        for (int i = info.n_rels; i >= 1; i--) {
            Bits<1> s;
            s.w[0] = 1ull << i;
            emit(s, Bits<1>());
            csg_rec(s, lower_inc(s), Bits<1>());
        }
    }
};

/* cpu_dpccp: the sequential framework with try_swap, i.e. both
 * orientations of join(memo[R], memo[L]) (gpuqo_cpu_sequential.cu:34-58). */
static QueryTree<1> *cpu_dpccp(const Info<1> &info, Counters &c) {
    Memo<1> memo;
    init_memo(info, memo);
    auto join = [&](const Bits<1> &l, const Bits<1> &r) {
        JR<1> &L = *memo.find(l)->second, &R = *memo.find(r)->second;
        c.checks++;
        JR<1> *o;
        if (do_join(R, L, info, memo, &o, c)) c.sets++;
        if (do_join(L, R, info, memo, &o, c)) c.sets++;
    };
    DPccpEnum<decltype(join)> e{info, join};
    e.run();
    return finish_memo(info, memo);
}

/* ------------------------------------------------------------------------
 * dpe_dpccp: DPccp feeding the DPE framework (gpuqo_cpu_dpe.cu:55-417,
 * gpuqo_dependency_buffer.cu:30-129).  The enumerating thread appends join
 * jobs to the "next" dependency buffer: jobs with the same target relation
 * form one ring, processed by one thread; targets are queued per size,
 * those whose operands are ready at the head.  A full buffer is swapped
 * with the "current" one once the gpuqo_dpe_n_threads - 1 workers (helped
 * by the enumerating thread) have drained it.  A job waits until both of
 * its operands have no pending ring.  As in the reference, a pair is
 * submitted twice in the same orientation (try_swap, :285-288), so this is
 * a timing baseline, not a parity oracle.
 * ---------------------------------------------------------------------- */
struct DpeRel {
    JR<1> r;
    std::atomic<int> num_entry{0};
};

struct DepNode {
    DpeRel *target, *left, *right;
    DepNode *next_join, *prev_join, *next_rel, *prev_rel;
};

struct DepBuf {
    int n_rels, capacity;
    std::vector<DepNode> nodes;
    int next_free = 0;
    std::vector<DepNode *> queues;
    std::atomic<DepNode *> unified{nullptr};
    std::unordered_map<uint64_t, DepNode *> lookup;
    DepBuf(int n, int cap) : n_rels(n), capacity(cap), nodes((size_t)cap), queues((size_t)n + 2, nullptr) {}
    bool full() const { return next_free >= capacity; }
    bool empty() const { return unified.load() == nullptr; }
    void clear() {
        next_free = 0;
        std::fill(queues.begin(), queues.end(), nullptr);
        lookup.clear();
        unified = nullptr;
    }
    void push(DpeRel *t, DpeRel *l, DpeRel *r) {               /* :30-95 */
        const int index = t->r.id.size();
        DepNode *nn = &nodes[(size_t)next_free++];
        nn->target = t; nn->left = l; nn->right = r;
        nn->next_rel = nn->prev_rel = nn->next_join = nn->prev_join = nn;
        const bool ready = l->num_entry.load() == 0 && r->num_entry.load() == 0;
        auto it = lookup.find(t->r.id.w[0]);
        if (queues[(size_t)index] == nullptr) {
            queues[(size_t)index] = nn;
            lookup.emplace(t->r.id.w[0], nn);
            t->num_entry++;
        } else if (it == lookup.end()) {
            DepNode *h = queues[(size_t)index];
            nn->next_rel = h;
            nn->prev_rel = h->prev_rel;
            h->prev_rel->next_rel = nn;
            h->prev_rel = nn;
            if (ready) queues[(size_t)index] = nn;
            lookup.emplace(t->r.id.w[0], nn);
            t->num_entry++;
        } else {
            DepNode *node = it->second;
            nn->next_join = node;
            nn->prev_join = node->prev_join;
            node->prev_join->next_join = nn;
            node->prev_join = nn;
            if (ready) {                                     /* the ready job represents the ring */
                if (node != node->next_rel) {
                    node->prev_rel->next_rel = nn;
                    node->next_rel->prev_rel = nn;
                    nn->prev_rel = node->prev_rel;
                    nn->next_rel = node->next_rel;
                }
                if (queues[(size_t)index] == node) queues[(size_t)index] = nn;
                it->second = nn;
            }
        }
    }
    void unify() {                                          /* :97-117 */
        DepNode *node = nullptr;
        for (size_t i = 0; i < queues.size(); i++) {
            if (!queues[i]) continue;
            if (node) {
                node->prev_rel->next_rel = queues[i];
                node = queues[i];
            } else {
                node = queues[i];
                unified = node;
            }
        }
        if (node) node->prev_rel->next_rel = nullptr;
    }
    DepNode *pop() {                                        /* :119-127 */
        DepNode *n = unified.load();
        while (n && !unified.compare_exchange_strong(n, n->next_rel)) {
        }
        return n;
    }
};

static void dpe_process(DepBuf *buf, const Info<1> &info, std::atomic<uint64_t> &joins) {   /* :60-93 */
    DepNode *job;
    while ((job = buf->pop()) != nullptr) {
        DpeRel *t = job->target;
        DepNode *first = job;
        do {
            while (job->left->num_entry.load() != 0) {
            }
            while (job->right->num_entry.load() != 0) {
            }
            const JR<1> &L = job->left->r, &R = job->right->r;
            JoinRel v = make_join(L.id, L.v, R.id, R.v, info);
            joins++;
            if (v.cost.total < t->r.v.cost.total) {
                t->r.v = v;
                t->r.left_id = L.id;
                t->r.right_id = R.id;
                t->r.lp = const_cast<JR<1> *>(&L);
                t->r.rp = const_cast<JR<1> *>(&R);
            }
        } while ((job = job->next_join) != first);
        t->num_entry--;
    }
}

static QueryTree<1> *dpe_dpccp(const Info<1> &info, int n_threads, int pairs_per_depbuf, Counters &c) {
    n_threads = std::max(1, n_threads);
    pairs_per_depbuf = std::max(2, pairs_per_depbuf);
    std::unordered_map<uint64_t, DpeRel *> memo;
    for (int i = 0; i < info.n_rels; i++) {                 /* :371-386 */
        DpeRel *d = new DpeRel();
        d->r.id = info.base_rels[i].id;
        d->r.lp = d->r.rp = nullptr;
        d->r.v.cost = cost_baserel(info.base_rels[i], info.params);
        d->r.v.width = info.base_rels[i].width;
        d->r.v.rows = info.base_rels[i].rows;
        d->r.edges = info.edge_table[i];
        memo.emplace(d->r.id.w[0], d);
    }
    DepBuf b0(info.n_rels, pairs_per_depbuf), b1(info.n_rels, pairs_per_depbuf);
    DepBuf *cur = &b0, *next = &b1;
    std::mutex mu;
    std::condition_variable avail, all_waiting;
    int n_waiting = 0;
    uint64_t generation = 0;
    bool finish = false;
    std::atomic<uint64_t> joins{0};
    auto worker = [&]() {                                   /* :95-140 */
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                while (generation == seen && !finish) {
                    n_waiting++;
                    if (n_waiting == n_threads - 1) all_waiting.notify_one();
                    avail.wait(lk, [&] { return generation != seen || finish; });
                }
                if (finish) break;
                seen = generation;
            }
            dpe_process(cur, info, joins);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads - 1; t++) th.emplace_back(worker);
    auto wait_and_swap = [&]() {                            /* :205-228 */
        next->unify();
        dpe_process(cur, info, joins);
        std::unique_lock<std::mutex> lk(mu);
        all_waiting.wait(lk, [&] { return n_waiting >= n_threads - 1; });
        std::swap(cur, next);
        n_waiting = 0;
        generation++;
        avail.notify_all();
        lk.unlock();
        next->clear();
    };
    auto submit = [&](DpeRel *l, DpeRel *r) {               /* :167-200 */
        Bits<1> id = l->r.id | r->r.id;
        DpeRel *t;
        auto it = memo.find(id.w[0]);
        if (it != memo.end()) {
            t = it->second;
        } else {
            t = new DpeRel();
            t->r.id = id;
            t->r.left_id = l->r.id; t->r.right_id = r->r.id;
            t->r.lp = &l->r; t->r.rp = &r->r;
            t->r.edges = l->r.edges | r->r.edges;
            t->r.v.cost.total = INFF_;
            t->r.v.cost.startup = INFF_;
            t->r.v.rows = INFF_;
            t->r.v.width = 0;
            memo.emplace(id.w[0], t);
            c.sets++;
        }
        next->push(t, l, r);
        if (next->full()) wait_and_swap();
    };
    auto join = [&](const Bits<1> &lid, const Bits<1> &rid) {   /* join(level, true, memo[R], memo[L]) */
        DpeRel *L = memo.find(rid.w[0])->second, *R = memo.find(lid.w[0])->second;
        c.checks++;
        submit(L, R);
        submit(L, R);                                       /* same orientation again (:285-288) */
    };
    DPccpEnum<decltype(join)> e{info, join};
    e.run();
    wait_and_swap();
    dpe_process(cur, info, joins);
    {
        std::lock_guard<std::mutex> g(mu);
        finish = true;
        avail.notify_all();
    }
    for (auto &t : th) t.join();
    c.joins += joins.load();
    Bits<1> full;
    for (int i = 0; i < info.n_rels; i++) full |= info.base_rels[i].id;
    QueryTree<1> *out = nullptr;
    auto f = memo.find(full.w[0]);
    if (f != memo.end()) build_query_tree(&f->second->r, &out);
    for (auto &kv : memo) delete kv.second;
    return out;
}

struct RunCtx {
    int algorithm;
    bool bicc;
    int idp_type, idp_n_iters;
    int n_threads, chunk, pairs_per_depbuf;
    Counters c;
    int dp_calls = 0;
    bool connected_only = false;     /* DPsub / MPDP over connected_enumerate_sets */
};

template <int W>
static QueryTree<W> *run_switch(RunCtx &ctx, const Info<W> &info);

template <>
QueryTree<4> *run_switch<4>(RunCtx &ctx, const Info<4> &info) {
    if (ctx.algorithm == GQ_ALGO_CPU_GOO) return cpu_goo(info);
    return nullptr;   /* gpuqo_main_internal.cu:61-79: DP needs < 64 relations */
}

template <>
QueryTree<1> *run_switch<1>(RunCtx &ctx, const Info<1> &info) {     /* gpuqo_main_internal.cu:19-59 */
    ctx.dp_calls++;
    switch (ctx.algorithm) {
    case GQ_ALGO_DPSUB:
        return ctx.bicc ? cpu_dpsub_bicc(info, ctx.c, ctx.connected_only)
                        : cpu_dpsub(info, ctx.c, ctx.connected_only);
    case GQ_ALGO_CPU_DPSUB:
        return cpu_dpsub(info, ctx.c, ctx.connected_only);
    case GQ_ALGO_CPU_DPSUB_BICC:
        return cpu_dpsub_bicc(info, ctx.c, ctx.connected_only);
    case GQ_ALGO_DPSIZE:
    case GQ_ALGO_CPU_DPSIZE:
        return cpu_dpsize(info, ctx.c);
    case GQ_ALGO_CPU_DPSUB_PARALLEL:
        return parallel_cpu_dpsub(info, false, ctx.n_threads, ctx.chunk, ctx.c);
    case GQ_ALGO_CPU_DPSUB_BICC_PARALLEL:
        return parallel_cpu_dpsub(info, true, ctx.n_threads, ctx.chunk, ctx.c);
    case GQ_ALGO_CPU_GOO:
        return cpu_goo(info);
    case GQ_ALGO_CPU_DPCCP:
        return cpu_dpccp(info, ctx.c);
    case GQ_ALGO_DPE_DPCCP:
        return dpe_dpccp(info, ctx.n_threads, ctx.pairs_per_depbuf, ctx.c);
    default:
        return nullptr;
    }
}

/* IDP1: gpuqo_idp.cu:38-80 */
template <int W>
static QueryTree<W> *run_idp1(RunCtx &ctx, Info<W> &info);

template <int WO, int WI>
static QueryTree<WO> *idp1_next(RunCtx &ctx, const Info<WO> &info, std::list<RemapEl<WO>> &lst) {
    Remapper<WO, WI> rm(lst);
    Info<WI> *ni = rm.remapPlannerInfo(info);
    QueryTree<WI> *qt = run_idp1(ctx, *ni);
    QueryTree<WO> *o = rm.remapQueryTree(qt);
    delete ni;
    return o;
}

template <int W>
static QueryTree<W> *run_idp1(RunCtx &ctx, Info<W> &info) {
    info.n_iters = std::min(info.n_rels, ctx.idp_n_iters);
    QueryTree<W> *qt = run_switch(ctx, info);
    if (!qt || info.n_iters == info.n_rels) return qt;
    std::list<RemapEl<W>> lst;
    lst.push_back(RemapEl<W>{qt->id, 0, qt});
    int j = 1;
    for (int i = 0; i < info.n_rels; i++)
        if (!info.base_rels[i].id.isSubset(qt->id)) lst.push_back(RemapEl<W>{info.base_rels[i].id, j++, nullptr});
    if (W == 1 || lst.size() < 64) return idp1_next<W, 1>(ctx, info, lst);
    return idp1_next<W, 4>(ctx, info, lst);
}

/* IDP2: gpuqo_idp.cu:95-297 */
template <int W>
static QueryTree<W> *find_most_expensive_subtree(QueryTree<W> *qt, int max_size) {
    if (qt->id.size() <= max_size) return qt;
    QueryTree<W> *l = find_most_expensive_subtree(qt->left, max_size);
    QueryTree<W> *r = find_most_expensive_subtree(qt->right, max_size);
    if (l->id.size() == 1) return r;
    else if (r->id.size() == 1) return l;
    else if (l->cost.total > r->cost.total) return l;
    else if (l->cost.total < r->cost.total) return r;
    else if (l->id.size() < r->id.size()) return l;
    else return r;
}

template <int WO, int WI>
static QueryTree<WO> *sub_dp(RunCtx &ctx, const Info<WO> &info, std::list<RemapEl<WO>> &lst) {
    /* gpuqo_run_idp2_dp :122-141, gpuqo_run_dpdp_union_dp union.cu:261-280 */
    Remapper<WO, WI> rm(lst);
    Info<WI> *ni = rm.remapPlannerInfo(info);
    ni->n_iters = ni->n_rels;
    QueryTree<WI> *qt = run_switch(ctx, *ni);
    QueryTree<WO> *o = rm.remapQueryTree(qt);
    freeQueryTree(qt);
    delete ni;
    return o;
}

template <int WO, int WI>
static QueryTree<WO> *idp2_rec(RunCtx &ctx, QueryTree<WO> *goo_qt, const Info<WO> &info,
                               std::list<RemapEl<WO>> &lst, int n_iters) {
    Remapper<WO, WI> rm(lst);
    Info<WI> *ni = rm.remapPlannerInfo(info);
    QueryTree<WI> *ngoo = rm.remapQueryTreeFwd(goo_qt);
    ni->n_iters = std::min(ni->n_rels, n_iters);
    QueryTree<WI> *maxqt = find_most_expensive_subtree(ngoo, ni->n_iters);
    Bits<WI> reop = maxqt->id;
    std::list<RemapEl<WI>> rl;
    int i = 0;
    while (!reop.empty()) {
        RemapEl<WI> el; el.from_relid = reop.lowest(); el.to_idx = i++; el.qt = nullptr;
        rl.push_back(el);
        reop -= el.from_relid;
    }
    QueryTree<WI> *reopt = (WI == 1 || rl.size() < 64) ? sub_dp<WI, 1>(ctx, *ni, rl) : nullptr;
    QueryTree<WI> *res;
    if (!reopt) res = nullptr;
    else if (ni->n_iters == ni->n_rels) res = reopt;
    else {
        std::list<RemapEl<WI>> nl;
        nl.push_back(RemapEl<WI>{reopt->id, 0, reopt});
        int j = 1;
        for (int k = 0; k < ni->n_rels; k++)
            if (!ni->base_rels[k].id.isSubset(reopt->id)) nl.push_back(RemapEl<WI>{ni->base_rels[k].id, j++, nullptr});
        if (WI == 1 || nl.size() < 64) res = idp2_rec<WI, 1>(ctx, ngoo, *ni, nl, n_iters);
        else res = idp2_rec<WI, 4>(ctx, ngoo, *ni, nl, n_iters);
    }
    QueryTree<WO> *o = rm.remapQueryTree(res);
    freeQueryTree(res);
    freeQueryTree(ngoo);
    delete ni;
    return o;
}

template <int W>
static QueryTree<W> *run_idp2(RunCtx &ctx, Info<W> &info) {
    QueryTree<W> *goo = cpu_goo(info);
    if (!goo) return nullptr;
    std::list<RemapEl<W>> lst = identity_list(info);
    QueryTree<W> *o = idp2_rec<W, W>(ctx, goo, info, lst, ctx.idp_n_iters);
    freeQueryTree(goo);
    return o;
}

/* UnionDP: gpuqo_dpdp_union.cu:35-472 */
template <int W>
struct GraphEdge {
    Bits<W> left, right;
    int left_size, right_size, total_size;
    float rows, selectivity;
    PathCost cost;
    int width;
    float weight;
};

template <int W>
struct CompareEdges {                 /* union.cu:147-155 */
    bool operator()(const GraphEdge<W> *l, const GraphEdge<W> *r) const {
        if (l->total_size == r->total_size) return l->weight > r->weight;
        return l->total_size > r->total_size;
    }
};

template <int W>
struct DisjointSet {                  /* union.cu:35-109 */
    std::unordered_map<Bits<W>, Bits<W>, BitsHash<W>> parent, csg;
    std::unordered_map<Bits<W>, int, BitsHash<W>> size;
    void makeSet(const Info<W> &info) {
        for (int i = 0; i < info.n_rels; i++) {
            Bits<W> id = info.base_rels[i].id;
            parent[id] = id; csg[id] = id; size[id] = 1;
        }
    }
    Bits<W> Find(const Bits<W> &x) {
        if (parent[x] != x) parent[x] = Find(parent[x]);
        return parent[x];
    }
    int getSize(const Bits<W> &x) { return size[Find(x)]; }
    Bits<W> getCsg(const Bits<W> &x) { return csg[Find(x)]; }
    void Union(const GraphEdge<W> *e) {
        Bits<W> x = Find(e->left), y = Find(e->right);
        if (x == y) return;
        if (size[x] > size[y]) { parent[y] = x; size[x] += size[y]; csg[x] |= csg[y]; }
        else if (size[x] < size[y]) { parent[x] = y; size[y] += size[x]; csg[y] |= csg[x]; }
        else { parent[x] = y; size[y] *= 2; csg[y] |= csg[x]; }
    }
};

template <int W>
static GraphEdge<W> *createGraphEdge(int li, int ri, const Info<W> &info) {   /* union.cu:163-210 */
    GraphEdge<W> *e = new GraphEdge<W>();
    e->left = info.base_rels[li].id;
    e->right = info.base_rels[ri].id;
    float lr = info.base_rels[li].rows, rr = info.base_rels[ri].rows;
    e->selectivity = estimate_join_selectivity(e->left, e->right, info);
    e->rows = e->selectivity * lr * rr;
    e->left_size = 1; e->right_size = 1; e->total_size = 2;
    JoinRel L, R;
    L.cost = cost_baserel(info.base_rels[li], info.params); L.width = info.base_rels[li].width; L.rows = info.base_rels[li].rows;
    R.cost = cost_baserel(info.base_rels[ri], info.params); R.width = info.base_rels[ri].width; R.rows = info.base_rels[ri].rows;
    e->cost = calc_join_cost(e->left, L, e->right, R, e->rows, info);
    e->width = L.width + R.width;
    e->weight = e->rows;
    return e;
}

template <int W>
static void fillEdgeQueue(std::vector<GraphEdge<W> *> &all,
                          std::priority_queue<GraphEdge<W> *, std::vector<GraphEdge<W> *>, CompareEdges<W>> &q,
                          const Info<W> &info) {                                /* union.cu:213-259 */
    std::queue<int> bq;
    int bfs_idx = 0;
    bq.push(0);
    Bits<W> seen2 = Bits<W>::nth(1);
    while (!bq.empty() && bfs_idx < info.n_rels) {
        int u = bq.front(); bq.pop();
        Bits<W> edges = info.edge_table[u];
        bfs_idx++;
        while (!edges.empty()) {
            int next = edges.lowestPos();
            if (!seen2.isSet(next)) {
                bq.push(next - 1);
                GraphEdge<W> *e = createGraphEdge(u, next - 1, info);
                q.push(e);
                all.push_back(e);
            }
            edges.unset(next);
        }
        seen2 |= Bits<W>::nth(u + 1);
    }
}

template <int WO, int WI>
static QueryTree<WO> *union_rec(RunCtx &ctx, const Info<WO> &info, std::list<RemapEl<WO>> &lst, int n_iters) {
    Remapper<WO, WI> rm(lst);
    Info<WI> *ni = rm.remapPlannerInfo(info);
    ni->n_iters = std::min(ni->n_rels, n_iters);
    if (ni->n_rels == ni->n_iters) {                                            /* :295-303 */
        QueryTree<WO> *o = (WI == 1) ? sub_dp<WO, WI>(ctx, info, lst) : nullptr;
        delete ni;
        return o;
    }
    std::priority_queue<GraphEdge<WI> *, std::vector<GraphEdge<WI> *>, CompareEdges<WI>> q;
    std::vector<GraphEdge<WI> *> all;
    fillEdgeQueue(all, q, *ni);
    DisjointSet<WI> ds;
    ds.makeSet(*ni);
    const int upper_threshold = 16;                                             /* :320 */
    while (!q.empty()) {                                                        /* :324-344 */
        GraphEdge<WI> *e = q.top(); q.pop();
        if (ds.Find(e->left) != ds.Find(e->right)) {
            if (e->total_size != ds.getSize(e->left) + ds.getSize(e->right)) {
                e->left_size = ds.getSize(e->left);
                e->right_size = ds.getSize(e->right);
                e->total_size = e->left_size + e->right_size;
                q.push(e);
            } else if (e->total_size < upper_threshold) {
                ds.Union(e);
            }
        }
    }
    for (auto *e : all) delete e;
    std::vector<Bits<WI>> subgraphs;                                            /* :357-367 */
    Bits<WI> seen;
    for (int i = 0; i < ni->n_rels; i++) {
        Bits<WI> g = ds.getCsg(ni->base_rels[i].id);
        if (!g.isSubset(seen)) subgraphs.push_back(g);
        seen |= g;
    }
    std::list<RemapEl<WI>> next;
    bool fail = false;
    for (size_t i = 0; i < subgraphs.size(); i++) {                             /* :386-419 */
        std::list<RemapEl<WI>> rl;
        int j = 0;
        Bits<WI> reop = subgraphs[i];
        while (!reop.empty()) {
            RemapEl<WI> el; el.from_relid = reop.lowest(); el.to_idx = j++; el.qt = nullptr;
            rl.push_back(el);
            reop -= el.from_relid;
        }
        QueryTree<WI> *r = (rl.size() < 64) ? sub_dp<WI, 1>(ctx, *ni, rl) : nullptr;
        if (!r) { fail = true; break; }
        next.push_back(RemapEl<WI>{r->id, (int)i, r});
    }
    QueryTree<WI> *res = nullptr;
    if (!fail) {
        if (WI == 1 || next.size() < 64) res = union_rec<WI, 1>(ctx, *ni, next, n_iters);
        else res = union_rec<WI, 4>(ctx, *ni, next, n_iters);
    }
    QueryTree<WO> *o = rm.remapQueryTree(res);
    freeQueryTree(res);
    delete ni;
    return o;
}

template <int W>
static QueryTree<W> *run_union(RunCtx &ctx, Info<W> &info) {
    std::list<RemapEl<W>> lst = identity_list(info);
    return union_rec<W, W>(ctx, info, lst, ctx.idp_n_iters);
}

/* __gpuqo_run: gpuqo_main_internal.cu:81-147 (the spanning tree, :94-100, is taken in convert) */
template <int W>
static QueryTree<W> *gpuqo_run_t(RunCtx &ctx, Info<W> &info) {
    std::list<RemapEl<W>> bfs = bfs_remap_list(info);
    Remapper<W, W> rm(bfs);
    Info<W> *ri = rm.remapPlannerInfo(info);
    QueryTree<W> *qt;
    if (ctx.idp_n_iters <= 1 || ctx.idp_n_iters >= info.n_rels) qt = run_switch(ctx, *ri);
    else if (ctx.idp_type == GQ_IDP1) qt = run_idp1(ctx, *ri);
    else if (ctx.idp_type == GQ_IDP2) qt = run_idp2(ctx, *ri);
    else qt = run_union(ctx, *ri);
    QueryTree<W> *o = rm.remapQueryTree(qt);
    freeQueryTree(qt);
    delete ri;
    return o;
}

/* ------------------------------------------------------------------------
 * C ABI glue: conversion from gq_query (0-based bits) and to gq_plan_node.
 * setParams (gpuqo_planner_info.cu:113-131) narrows the GUC doubles to float.
 * ---------------------------------------------------------------------- */
static Params make_params(const gq_config *c) {
    Params p;
    p.effective_cache_size = (float)c->effective_cache_size;
    p.random_page_cost = (float)c->random_page_cost;
    p.seq_page_cost = (float)c->seq_page_cost;
    p.cpu_tuple_cost = (float)c->cpu_tuple_cost;
    p.cpu_index_tuple_cost = (float)c->cpu_index_tuple_cost;
    p.cpu_operator_cost = (float)c->cpu_operator_cost;
    p.disable_cost = (float)c->disable_cost;
    p.enable_seqscan = c->enable_seqscan; p.enable_indexscan = c->enable_indexscan;
    p.enable_tidscan = c->enable_tidscan; p.enable_sort = c->enable_sort;
    p.enable_hashagg = c->enable_hashagg; p.enable_nestloop = c->enable_nestloop;
    p.enable_mergejoin = c->enable_mergejoin; p.enable_hashjoin = c->enable_hashjoin;
    p.work_mem = c->work_mem;
    p.cost_function = c->cost_function;
    return p;
}

template <int W>
static Bits<W> from_words(const uint64_t *w, int nw) {   /* relation i -> bit i+1 */
    Bits<W> b;
    for (int i = 0; i < nw * 64; i++)
        if ((w[i >> 6] >> (i & 63)) & 1ULL) b.set(i + 1);
    return b;
}

/* minimumSpanningTree: gpuqo_spanning_tree.cu:17-68.  Grows S from base
 * relation 0; each round scans i in S ascending and j in edge_table[i] - S
 * ascending, keeps the first strict minimum of estimate_join_rows of the two
 * base relations against FLT_MAX, and adds that edge to an empty edge table.
 * estimate_join_rows clamps a NaN product to 1 row, so an edge with a NaN
 * selectivity weighs 1; only +inf rows are never taken.
 * Only the edge table is replaced.  Where no edge passes the test (the
 * reference then reads an unset argmin) the tree built so far is kept. */
template <int W>
static void minimum_spanning_tree(Info<W> &info) {
    const int n = info.n_rels;
    Bits<W> S = info.base_rels[0].id, all;
    for (int i = 0; i < n; i++) all |= info.base_rels[i].id;
    std::vector<Bits<W>> out(n);
    while (S != all) {
        float best = FLT_MAX;
        int ai = -1, aj = -1;
        for (int i = 0; i < n; i++) {
            if (!S.isSet(i + 1)) continue;
            Bits<W> edges = info.edge_table[i] - S;
            for (int j = 0; j < n; j++) {
                if (!edges.isSet(j + 1)) continue;
                const float rows = estimate_join_rows(info.base_rels[i].id, info.base_rels[i].rows,
                                                      info.base_rels[j].id, info.base_rels[j].rows, info);
                if (rows < best) { best = rows; ai = i; aj = j; }
            }
        }
        if (ai < 0) break;
        S.set(aj + 1);
        out[ai].set(aj + 1);
        out[aj].set(ai + 1);
    }
    info.edge_table.swap(out);
}

template <int W>
static Info<W> *convert(const gq_query *q, const gq_config *cfg) {   /* convertGpuqoPlannerInfo :133-209 */
    Info<W> *info = new Info<W>();
    info->n_rels = q->n_rels;
    info->n_iters = q->n_rels;
    info->params = make_params(cfg);
    info->base_rels.resize(q->n_rels);
    info->edge_table.resize(q->n_rels);
    for (int i = 0; i < q->n_rels; i++) {
        info->edge_table[i] = from_words<W>(q->edge_table + (size_t)i * q->n_words, q->n_words);
        BaseRel<W> &b = info->base_rels[i];
        b.id = Bits<W>::nth(i + 1);
        b.rows = q->base_rels[i].rows; b.tuples = q->base_rels[i].tuples;
        b.width = q->base_rels[i].width;
        b.cost.startup = q->base_rels[i].cost.startup; b.cost.total = q->base_rels[i].cost.total;
        b.composite = false;
    }
    info->n_ec = q->n_eq_classes;
    int sel_off = 0;
    for (int e = 0; e < q->n_eq_classes; e++) {
        int b0 = q->ec_member_offsets[e], b1 = q->ec_member_offsets[e + 1];
        Bits<W> ec;
        for (int m = b0; m < b1; m++) ec.set(q->ec_members[m] + 1);
        info->ec_relids.push_back(ec);
        int s = b1 - b0;
        for (int j = 0; j < eqClassNSels(s); j++) info->ec_sels.push_back(q->ec_sels[sel_off + j]);
        sel_off += eqClassNSels(s);
        for (int m = b0; m < b1; m++) {
            info->ec_fks.push_back(from_words<W>(q->ec_fks + (size_t)m * q->n_words, q->n_words));
            info->ec_vars.push_back(q->ec_vars[m]);
        }
    }
    if (cfg->gpuqo_spanning_tree) minimum_spanning_tree(*info);
    return info;
}

template <int W>
static int emit_tree(const QueryTree<W> *qt, gq_plan_out *out) {   /* pre-order */
    if (!qt) return -1;
    if (out->n_nodes >= out->capacity) return -2;
    int me = out->n_nodes++;
    gq_plan_node &n = out->nodes[me];
    memset(n.relids, 0, sizeof(n.relids));
    for (int i = 1; i < W * 64; i++)
        if (qt->id.isSet(i)) n.relids[(i - 1) >> 6] |= 1ULL << ((i - 1) & 63);
    n.rows = qt->rows; n.cost.startup = qt->cost.startup; n.cost.total = qt->cost.total;
    n.width = qt->width;
    int l = emit_tree(qt->left, out);
    if (l == -2) return -2;
    int r = emit_tree(qt->right, out);
    if (r == -2) return -2;
    out->nodes[me].left = l;
    out->nodes[me].right = r;
    return me;
}

template <int W>
static int plan_t(const gq_query *q, const gq_config *cfg, gq_plan_out *out, bool connected_only = false) {
    RunCtx ctx;
    ctx.connected_only = connected_only;
    ctx.algorithm = cfg->gpuqo_algorithm;
    ctx.bicc = cfg->gpuqo_dpsub_bicc != 0;
    ctx.idp_type = cfg->gpuqo_idp_type;
    ctx.idp_n_iters = cfg->gpuqo_idp_n_iters;
    ctx.n_threads = cfg->gpuqo_dpe_n_threads > 0 ? cfg->gpuqo_dpe_n_threads : 1;
    ctx.chunk = cfg->gpuqo_cpu_dpsub_parallel_chunk_size > 0 ? cfg->gpuqo_cpu_dpsub_parallel_chunk_size : 64;
    ctx.pairs_per_depbuf = cfg->gpuqo_dpe_pairs_per_depbuf > 0 ? cfg->gpuqo_dpe_pairs_per_depbuf : 128;
    Info<W> *info = convert<W>(q, cfg);
    auto t0 = std::chrono::steady_clock::now();
    QueryTree<W> *qt = gpuqo_run_t(ctx, *info);
    auto t1 = std::chrono::steady_clock::now();
    delete info;
    out->n_nodes = 0;
    memset(&out->stats, 0, sizeof(out->stats));
    out->stats.sets = ctx.c.sets;
    out->stats.pairs = ctx.c.joins;
    out->stats.candidates = ctx.c.checks;
    out->stats.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    out->stats.n_dp_calls = ctx.dp_calls;
    if (!qt) return GQ_ERR_NO_PLAN;
    int r = emit_tree(qt, out);
    freeQueryTree(qt);
    return r == -2 ? GQ_ERR_CAPACITY : GQ_OK;
}

}  // namespace oracle

extern "C" {

/* Plan with the CPU restatement.  cfg->gpuqo_algorithm selects the
 * algorithm; GQ_ALGO_DPSUB maps to cpu_dpsub (bicc off) or cpu_dpsub_bicc
 * (bicc on) and GQ_ALGO_DPSIZE to cpu_dpsize -- the oracles the GPU must
 * match. */
int oracle_plan(const gq_query *q, const gq_config *cfg, gq_plan_out *out) {
    if (!q || !cfg || !out || q->n_rels < 1 || q->n_rels >= GQ_MAX_RELS) return GQ_ERR_INVALID;
    if (q->n_rels < 64) return oracle::plan_t<1>(q, cfg, out);
    return oracle::plan_t<4>(q, cfg, out);
}

/* oracle_plan with DPsub and MPDP (also inside IDP1 / IDP2 / UnionDP)
 * visiting the connected sets size by size (connected_enumerate_sets)
 * instead of all 2^n integers: the same plans, for DPs above 32 relations
 * whose 2^n pass would not finish. */
int oracle_plan_connected(const gq_query *q, const gq_config *cfg, gq_plan_out *out) {
    if (!q || !cfg || !out || q->n_rels < 1 || q->n_rels >= GQ_MAX_RELS) return GQ_ERR_INVALID;
    if (q->n_rels < 64) return oracle::plan_t<1>(q, cfg, out, true);
    return oracle::plan_t<4>(q, cfg, out, true);
}

/* Cost one join (outer = l, inner = r) of two already-costed relations with
 * the restated cost model, on the query as given (no relabelling). */
int oracle_join_cost(const gq_query *q, const gq_config *cfg, const uint64_t *l_words, const uint64_t *r_words,
                     const float *lvals, const float *rvals, float *out_vals) {
    /* vals = {rows, startup, total, width} */
    if (q->n_rels >= 64) return GQ_ERR_UNSUPPORTED;
    oracle::Info<1> *info = oracle::convert<1>(q, cfg);
    oracle::Bits<1> l = oracle::from_words<1>(l_words, q->n_words), r = oracle::from_words<1>(r_words, q->n_words);
    oracle::JoinRel L, R;
    L.rows = lvals[0]; L.cost.startup = lvals[1]; L.cost.total = lvals[2]; L.width = (int)lvals[3];
    R.rows = rvals[0]; R.cost.startup = rvals[1]; R.cost.total = rvals[2]; R.width = (int)rvals[3];
    oracle::JoinRel j = oracle::make_join(l, L, r, R, *info);
    out_vals[0] = j.rows; out_vals[1] = j.cost.startup; out_vals[2] = j.cost.total; out_vals[3] = (float)j.width;
    delete info;
    return GQ_OK;
}

/* A bounded sample of one DP for CPU-baseline timing: the algorithm of
 * cfg->gpuqo_algorithm runs its levels 2..max_levels only (the reference's
 * n_iters < n_rels: gpuqo_cpu_sequential.cu:93-113), once, on the
 * relabelled query; stats->pairs and stats->wall_ms cover that run. */
int oracle_dp_first_levels(const gq_query *q, const gq_config *cfg, int max_levels, gq_plan_stats *stats) {
    if (!q || !cfg || !stats || q->n_rels < 2 || q->n_rels >= 64 || max_levels < 2) return GQ_ERR_INVALID;
    oracle::RunCtx ctx;
    ctx.algorithm = cfg->gpuqo_algorithm;
    ctx.bicc = cfg->gpuqo_dpsub_bicc != 0;
    ctx.idp_type = cfg->gpuqo_idp_type;
    ctx.idp_n_iters = 0;
    ctx.n_threads = cfg->gpuqo_dpe_n_threads > 0 ? cfg->gpuqo_dpe_n_threads : 1;
    ctx.chunk = cfg->gpuqo_cpu_dpsub_parallel_chunk_size > 0 ? cfg->gpuqo_cpu_dpsub_parallel_chunk_size : 64;
    ctx.pairs_per_depbuf = cfg->gpuqo_dpe_pairs_per_depbuf > 0 ? cfg->gpuqo_dpe_pairs_per_depbuf : 128;
    oracle::Info<1> *info = oracle::convert<1>(q, cfg);
    std::list<oracle::RemapEl<1>> bfs = oracle::bfs_remap_list(*info);
    oracle::Remapper<1, 1> rm(bfs);
    oracle::Info<1> *ri = rm.remapPlannerInfo(*info);
    ri->n_iters = std::min(ri->n_rels, max_levels);
    auto t0 = std::chrono::steady_clock::now();
    oracle::QueryTree<1> *qt = oracle::run_switch(ctx, *ri);
    auto t1 = std::chrono::steady_clock::now();
    oracle::freeQueryTree(qt);
    delete ri;
    delete info;
    memset(stats, 0, sizeof(*stats));
    stats->sets = ctx.c.sets;
    stats->pairs = ctx.c.joins;
    stats->candidates = ctx.c.checks;
    stats->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    stats->n_dp_calls = 1;
    return GQ_OK;
}

/* One branch of the restated cost model, exposed for known-answer tests:
 *   which 0: calc_join_cost (gpuqo_cost.cuh:76-158) with the relations in
 *            composite_mask (bit i = relation i) marked composite;
 *   which 1: cost_nestloop (cost_postgres.cuh:110-275) and
 *   which 2: cost_hashjoin (:287-470) with CostExtra taken from extra_in =
 *            {inner_unique, indexed_join_quals, joinrows, outer_match_frac,
 *            match_count};
 *   which 3: estimate_hash_bucket_stats (:620-668), nbuckets = extra_in[0];
 *   which 4: is_inner_unique (gpuqo_cost.cuh:43-74);
 *   which 5: cost_baserel of l's lowest relation (cost_postgres.cuh:843-847,
 *            cost_simple.cuh:30-37, cost_cout.cuh:26-33).
 * Which 0-2 and 5 follow cfg->cost_function.
 * out = {startup, total} (0-2, 5), {mcvfreq, bucketsize} (3), {flag} (4). */
int oracle_cost_probe(const gq_query *q, const gq_config *cfg, int which, const uint64_t *l_words,
                      const uint64_t *r_words, const float *lvals, const float *rvals, const float *extra_in,
                      uint64_t composite_mask, float *out) {
    if (q->n_rels >= 64) return GQ_ERR_UNSUPPORTED;
    oracle::Info<1> *info = oracle::convert<1>(q, cfg);
    for (int i = 0; i < q->n_rels; i++) info->base_rels[i].composite = (composite_mask >> i) & 1ULL;
    oracle::Bits<1> l = oracle::from_words<1>(l_words, q->n_words), r = oracle::from_words<1>(r_words, q->n_words);
    oracle::JoinRel L, R;
    L.rows = lvals[0]; L.cost.startup = lvals[1]; L.cost.total = lvals[2]; L.width = (int)lvals[3];
    R.rows = rvals[0]; R.cost.startup = rvals[1]; R.cost.total = rvals[2]; R.width = (int)rvals[3];
    oracle::CostExtra ex;
    ex.inner_unique = extra_in[0] != 0.0f; ex.indexed_join_quals = extra_in[1] != 0.0f;
    ex.joinrows = extra_in[2]; ex.outer_match_frac = extra_in[3]; ex.match_count = extra_in[4];
    int st = GQ_OK;
    switch (which) {
    case 0: {
        oracle::PathCost c = oracle::calc_join_cost(l, L, r, R, extra_in[2], *info);
        out[0] = c.startup; out[1] = c.total;
        break;
    }
    case 1: { oracle::PathCost c = oracle::nestloop_of(l, L, r, R, ex, *info); out[0] = c.startup; out[1] = c.total; break; }
    case 2: { oracle::PathCost c = oracle::hashjoin_of(l, L, r, R, ex, *info); out[0] = c.startup; out[1] = c.total; break; }
    case 3: {
        oracle::BucketStats b = oracle::estimate_hash_bucket_stats(l, r, (int)extra_in[0], *info);
        out[0] = b.mcvfreq; out[1] = b.bucketsize;
        break;
    }
    case 4: out[0] = oracle::is_inner_unique(l, r, *info) ? 1.0f : 0.0f; break;
    case 5: {   /* cost_baserel of the lowest relation of l */
        oracle::PathCost c = oracle::cost_baserel(info->base_rels[l.lowestPos() - 1], info->params);
        out[0] = c.startup; out[1] = c.total;
        break;
    }
    default: st = GQ_ERR_INVALID;
    }
    delete info;
    return st;
}

/* __estimate_hash_bucketsize (cost_postgres.cuh:701-775) of one variable of
 * a base relation with the given rows and tuples, for known-answer tests. */
float oracle_hash_bucketsize(const gq_var_info *v, float rows, float tuples, int nbuckets) {
    oracle::BaseRel<1> b;
    b.rows = rows; b.tuples = tuples; b.width = 0; b.cost.startup = 0.0f; b.cost.total = 0.0f; b.composite = false;
    return oracle::estimate_hash_bucketsize<1>(*v, b, nbuckets);
}

/* ExecChooseHashTableSize, exposed for known-answer tests. */
void oracle_hash_table_size(float ntuples, int width, int work_mem, int *nbuckets, int *nbatches, int *nskew) {
    oracle::HtszRes r = oracle::exec_choose_hash_table_size(ntuples, width, work_mem);
    *nbuckets = r.numbuckets; *nbatches = r.numbatches; *nskew = r.num_skew_mcvs;
}

/* BFS relabelling order (new index -> old relation), for tests. */
int oracle_bfs_order(const gq_query *q, int *order) {
    if (q->n_rels >= 64) return GQ_ERR_UNSUPPORTED;
    gq_config cfg; memset(&cfg, 0, sizeof(cfg)); cfg.work_mem = 4096;
    oracle::Info<1> *info = oracle::convert<1>(q, &cfg);
    auto lst = oracle::bfs_remap_list(*info);
    int k = 0;
    for (auto &e : lst) order[k++] = e.from_relid.lowestPos() - 1;
    delete info;
    return k;
}

/* Blocks of a connected set in find_blocks emission order (0-based bits). */
int oracle_blocks(const gq_query *q, uint64_t set0, uint64_t *blocks_out, int cap) {
    if (q->n_rels >= 63) return -1;
    gq_config cfg; memset(&cfg, 0, sizeof(cfg)); cfg.work_mem = 4096;
    oracle::Info<1> *info = oracle::convert<1>(q, &cfg);
    std::vector<oracle::Bits<1>> blocks;
    oracle::find_blocks(oracle::Bits<1>(set0 << 1), info->edge_table, blocks);
    int k = 0;
    for (auto &b : blocks) { if (k < cap) blocks_out[k] = b.w[0] >> 1; k++; }
    delete info;
    return k;
}

/* Iteration order of the level-1 list of cpu_dpsize (libstdc++ unordered_map
 * over the base relations inserted in index order), cpu_dpsize.cu:43-47. */
int oracle_dpsize_level1_order(int n_rels, int *order) {
    std::unordered_map<oracle::Bits<1>, int, oracle::BitsHash<1>> m;
    for (int i = 0; i < n_rels; i++) m.insert(std::make_pair(oracle::Bits<1>::nth(i + 1), i));
    int k = 0;
    for (auto &p : m) order[k++] = p.second;
    return k;
}

}  // extern "C"
