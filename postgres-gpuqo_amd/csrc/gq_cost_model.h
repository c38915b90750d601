/*
 * gq_cost_model.h -- the Postgres-12 join cost model of the reference, in
 * float, written once for the GPU kernels and the host drivers.
 *
 * Follows src/backend/optimizer/gpuqo/gpuqo_cost.cuh:43-158,
 * gpuqo_cost_postgres.cuh:89-863 and gpuqo_row_estimation.cuh:26-158:
 * every float operation happens in the reference's order, so results are
 * bit-identical when compiled with -ffp-contract=off and IEEE division.
 *
 * Layout difference (not an arithmetic one): instead of testing every
 * equivalence class (EC) per join, the caller passes the bitmask of ECs that
 * have members on both sides, built from per-relation incidence masks, and
 * only those are visited, in ascending EC index.  Classes with no member on
 * one side contribute selectivity 1.0f (an exact multiply), no qual and no
 * bucket-size or index candidate in the reference, so skipping them changes
 * no bit.
 */
#ifndef GQ_COST_MODEL_H
#define GQ_COST_MODEL_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GQ_HD __host__ __device__ __forceinline__
#else
#define GQ_HD inline
#endif

#define GQ_MAX_ECW 8            /* EC masks: up to 512 equivalence classes */

namespace gq {

/* ---- relation-set operations for the set types used (u32 / u64 / SetN) -- */
GQ_HD int popc(uint32_t x) { return __builtin_popcount(x); }
GQ_HD int popc(uint64_t x) { return __builtin_popcountll(x); }
GQ_HD int ctz(uint32_t x) { return __builtin_ctz(x); }
GQ_HD int ctz(uint64_t x) { return __builtin_ctzll(x); }
GQ_HD bool empty(uint32_t x) { return x == 0; }
GQ_HD bool empty(uint64_t x) { return x == 0; }
GQ_HD bool bit(uint32_t x, int i) { return (x >> i) & 1u; }
GQ_HD bool bit(uint64_t x, int i) { return (x >> i) & 1ull; }
/* number of members of `y` below the lowest member of `x` */
GQ_HD int rank_below(uint32_t y, uint32_t x) { return popc(y & ((x & (0u - x)) - 1u)); }
GQ_HD int rank_below(uint64_t y, uint64_t x) {
    uint64_t low = x & ((uint64_t)0 - x);
    return __builtin_popcountll(y & (low - (uint64_t)1));
}
GQ_HD uint32_t lowbit(uint32_t x) { return x & (0u - x); }
GQ_HD uint64_t lowbit(uint64_t x) { return x & ((uint64_t)0 - x); }

/* Multi-word set for host drivers above 64 relations (bit i = relation i). */
template <int W>
struct SetN {
    uint64_t w[W];
    GQ_HD SetN() { for (int i = 0; i < W; i++) w[i] = 0; }
    GQ_HD static SetN nth(int n) { SetN s; s.w[n >> 6] = 1ull << (n & 63); return s; }
    GQ_HD SetN operator&(const SetN &o) const { SetN r; for (int i = 0; i < W; i++) r.w[i] = w[i] & o.w[i]; return r; }
    GQ_HD SetN operator|(const SetN &o) const { SetN r; for (int i = 0; i < W; i++) r.w[i] = w[i] | o.w[i]; return r; }
    GQ_HD SetN operator-(const SetN &o) const { SetN r; for (int i = 0; i < W; i++) r.w[i] = w[i] & ~o.w[i]; return r; }
    GQ_HD SetN &operator|=(const SetN &o) { for (int i = 0; i < W; i++) w[i] |= o.w[i]; return *this; }
    GQ_HD SetN &operator&=(const SetN &o) { for (int i = 0; i < W; i++) w[i] &= o.w[i]; return *this; }
    GQ_HD SetN &operator-=(const SetN &o) { for (int i = 0; i < W; i++) w[i] &= ~o.w[i]; return *this; }
    GQ_HD bool operator==(const SetN &o) const { for (int i = 0; i < W; i++) if (w[i] != o.w[i]) return false; return true; }
    GQ_HD bool operator!=(const SetN &o) const { return !(*this == o); }
    GQ_HD bool operator<(const SetN &o) const {
        for (int i = W - 1; i >= 0; i--) if (w[i] != o.w[i]) return w[i] < o.w[i];
        return false;
    }
    GQ_HD void set(int n) { w[n >> 6] |= 1ull << (n & 63); }
    GQ_HD void unset(int n) { w[n >> 6] &= ~(1ull << (n & 63)); }
};
template <int W> GQ_HD int popc(const SetN<W> &x) { int s = 0; for (int i = 0; i < W; i++) s += popc(x.w[i]); return s; }
template <int W> GQ_HD bool empty(const SetN<W> &x) { for (int i = 0; i < W; i++) if (x.w[i]) return false; return true; }
template <int W> GQ_HD int ctz(const SetN<W> &x) { for (int i = 0; i < W; i++) if (x.w[i]) return i * 64 + ctz(x.w[i]); return -1; }
template <int W> GQ_HD bool bit(const SetN<W> &x, int i) { return (x.w[i >> 6] >> (i & 63)) & 1ull; }
template <int W> GQ_HD SetN<W> lowbit(const SetN<W> &x) { SetN<W> r; int p = ctz(x); if (p >= 0) r.set(p); return r; }
template <int W> GQ_HD int rank_below(const SetN<W> &y, const SetN<W> &x) {
    int p = ctz(x), s = 0;
    for (int i = 0; i < W; i++) {
        uint64_t m;
        if (p >= (i + 1) * 64) m = ~0ull;
        else if (p > i * 64) m = (1ull << (p - i * 64)) - 1;
        else m = 0;
        s += popc(y.w[i] & m);
    }
    return s;
}

/* ---- query description seen by the cost model ------------------------- */
struct Params {             /* GpuqoPlannerInfoParams subset (float), planner_info.cuh:79-96 */
    float seq_page_cost;
    float cpu_tuple_cost;
    float cpu_operator_cost;
    float disable_cost;
    int work_mem;
    int enable_nestloop;
    int enable_hashjoin;
    int cost_fn;            /* COST_FN_*: the reference's build-time GPUQO_COST_FUNCTION */
};

/* Cost functions (gpuqo_cost.cuh:25-38): postgres (gpuqo_cost_postgres.cuh),
 * simple (gpuqo_cost_simple.cuh) and C_out (gpuqo_cost_cout.cuh). */
#define COST_FN_POSTGRES 0
#define COST_FN_SIMPLE 1
#define COST_FN_COUT 2

/* cost_baserel: the path cost of a base relation as a DP's memo starts
 * with it.  postgres (cost_postgres.cuh:843-847): the relation's own cost;
 * simple (cost_simple.cuh:30-37): startup 0, total 0.2f * tuples unless the
 * relation is composite (an IDP or UnionDP sub-plan: its total); C_out
 * (cost_cout.cuh:26-33): startup 0, total 0 unless composite. */
GQ_HD void base_cost(const Params &p, float tuples, bool composite, float *startup, float *total) {
    if (p.cost_fn == COST_FN_SIMPLE) {
        *startup = 0.0f;
        if (!composite) *total = 0.2f * tuples;
    } else if (p.cost_fn == COST_FN_COUT) {
        *startup = 0.0f;
        if (!composite) *total = 0.0f;
    }
}

/* Member flags */
#define GQ_MF_BUCKET_DEFAULT 1u   /* stadistinct resolves to <= 0: bucket size 0.1 */
#define GQ_MF_INDEX 2u            /* index available                               */
#define GQ_MF_UNIQUE 4u           /* index unique                                  */

/* Three records per two-member class (View::pair_rec): {sel_lo, sel_hi,
 * lower member as a float, 0}, then per member (lower, higher) the bucket
 * constants {ndistinct, 1/ndistinct, skew, mcvfreq}; a member whose bucket
 * size is the 0.1 default has {-inf, 0.1f, 1.0f, mcvfreq}, which gives
 * exactly 0.1f below. */
struct alignas(16) PairRec { float x, y, z, w; };

template <class S>
struct View {
    int n_rels, n_ec, ecw;
    Params p;
    const float *rel_rows;        /* [n_rels] base rows (bucket-size path)         */
    const float *rel_fksel;       /* [n_rels] float(1.0 / max(1.0f, tuples))      */
    S composite;                  /* relations that are composites (UnionDP)       */
    const uint64_t *inc;          /* [n_rels * ecw]: ECs containing the relation   */
    const S *ec_relids;           /* [n_ec]                                        */
    const int *ec_moff;           /* [n_ec] first member index                     */
    const int *ec_soff;           /* [n_ec] first selectivity index                */
    const uint8_t *ec_hasfk;      /* [n_ec] some member holds an FK                */
    const S *mem_fks;             /* [n_members]                                   */
    const float *mem_nd;          /* ndistinct after the rows/tuples adjustment    */
    const float *mem_inv_nd;      /* 1.0f / mem_nd                                 */
    const float *mem_skew;        /* mcvfreq / avgfreq where it applies, else 1.0f */
    const float *mem_mcv;
    const uint8_t *mem_flags;
    const float *mem_idx_rows, *mem_idx_startup, *mem_idx_total;
    const float *sels;
    const float *qualtab;         /* qualtab[k] = cpu_operator_cost added k times  */
    /* pairs != 0: every class has exactly two members; class e's lower
     * member is relation ec_lo[e], and its selectivity is sel_lo[e] when
     * that member is on the outer side of a join, sel_hi[e] otherwise */
    int pairs;
    const PairRec *pair_rec;      /* [3 * n_ec] when pairs                         */
};

struct Rel {                /* value part of JoinRelation (planner_info.cuh:22-28) */
    float rows, startup, total;
    int width;
    uint32_t hs;             /* 0, or hash_pack(): the relation's hash table size as an inner */
};

GQ_HD int eq_class_index(int l, int r, int size) {    /* gpuqo_common.h:71-79 */
    if (l > r) { int t = l; l = r; r = t; }
    return l * size - l * (l + 1) / 2 + (r - l - 1);
}

/* float -> int as the reference's x86 host build converts it (cvttss2si):
 * values outside int32 range and NaN become INT_MIN.  Used where the
 * reference passes a float to an int (cost_postgres.cuh:370, :573, :615). */
GQ_HD int f2i_x86(float f) {
    if (f >= -2147483648.0f && f < 2147483648.0f) return (int)f;
    return (int)0x80000000u;
}

GQ_HD int floor_log2_l(long num) { return 63 - __builtin_clzll((unsigned long long)num); }
GQ_HD int ceil_log2_l(long num) {
    int r = floor_log2_l(num);
    return (1L << r) == num ? r : r + 1;
}

GQ_HD float clamp_row_est(float nrows) {               /* cost_postgres.cuh:850-863 */
    return nrows <= 1.0f ? 1.0f : rintf(nrows);
}
GQ_HD unsigned long maxalign8(long len) { return ((unsigned long)len + 7ul) & ~7ul; }
/* The C ABI admits widths of 0 .. 2^31 - 1 only (base widths are checked
 * non-negative and every sum of them below 2^31), so maxalign8(width) + 24
 * is below 2^32 and its u32 conversion rounds the same value as the
 * reference's conversion of the 64-bit one (a 64-bit integer conversion is
 * a long instruction sequence on the GPU). */
GQ_HD float relation_byte_size(float tuples, int width) {   /* :822-826 */
    return tuples * (float)(uint32_t)(maxalign8(width) + 24ul);
}

GQ_HD uint32_t f_bits(float x) { uint32_t b; memcpy(&b, &x, 4); return b; }
GQ_HD float bits_f(uint32_t b) { float x; memcpy(&x, &b, 4); return x; }

/* 1 / x for x = +-2^e, a normal float whose reciprocal is normal too, else
 * 0.  Then the reciprocal is exact, so x * (1 / x) is never needed: a / x
 * and a * recip_pow2(x) are the same correctly rounded value of the same
 * real quotient, for every a (subnormal, infinite and NaN a included). */
GQ_HD float recip_pow2(float x) {
    const uint32_t b = f_bits(x), e = (b >> 23) & 255u;
    if ((b & 0x7fffffu) || e == 0u || e > 253u) return 0.0f;
    return bits_f((b & 0x80000000u) | ((254u - e) << 23));
}
/* a / x through the exact reciprocal when x is a power of two */
GQ_HD float div_pow2(float a, float x, float rx) { return rx != 0.0f ? a * rx : a / x; }
GQ_HD float page_size(float tuples, int width) {            /* :832-840 */
    return ceilf(relation_byte_size(tuples, width) / 8192.0f);
}

struct HashSize { int nbuckets, nbatches; };

/* a / b for long operands; the same quotient through a 32-bit unsigned
 * divide when both are in [0, 2^32) (a 64-bit divide is a long software
 * sequence on the GPU). */
GQ_HD long div_l(long a, long b) {
    if (a >= 0 && b > 0 && a <= 0xffffffffL && b <= 0xffffffffL) return (long)((uint32_t)a / (uint32_t)b);
    return a / b;
}

/* ExecChooseHashTableSize with NTUP_PER_BUCKET 1, SKEW_WORK_MEM_PERCENT 2,
 * SKEW_OVERHEAD 84, TUPLE_OVERHEAD 32, MaxAllocSize 0x3fffffff and 8-byte
 * pointers (cost_postgres.cuh:472-618).  min/max on float are fminf/fmaxf. */
GQ_HD HashSize hash_table_size(float ntuples, int tupwidth, int work_mem) {
    HashSize res;
    res.nbatches = 1;
    if (ntuples <= 0.0f) ntuples = 1000.0f;
    int tupsize = (int)(32ul + maxalign8(tupwidth));
    float inner_rel_bytes = ntuples * (float)tupsize;
    long hash_table_bytes = (long)work_mem * 1024L;
    float space_allowed = (float)hash_table_bytes;
    long skew_table_bytes = hash_table_bytes * 2 / 100;
    /* (int)(skew_table_bytes / (tupsize + 84)) > 0.  For tupwidth >= 0 the
     * divisor is at least 116 and, work_mem being an int, the dividend below
     * 2^36, so the quotient fits an int and the test is a comparison. */
    const bool skew = tupwidth >= 0 ? skew_table_bytes >= (long)(tupsize + 84)
                                    : (int)div_l(skew_table_bytes, (long)(tupsize + 84)) > 0;
    if (skew) hash_table_bytes -= skew_table_bytes;
    long max_pointers = (long)(space_allowed / 8.0f);
    const long alloc_ptrs = 0x3fffffffL / 8;
    if (alloc_ptrs < max_pointers) max_pointers = alloc_ptrs;
    long mppow2 = 1L << ceil_log2_l(max_pointers);
    if (max_pointers != mppow2) max_pointers = mppow2 / 2;
    if (max_pointers > 0x7fffffffL / 2) max_pointers = 0x7fffffffL / 2;
    float dbuckets = ceilf(ntuples);
    dbuckets = fminf(dbuckets, (float)max_pointers);
    int nb = f2i_x86(dbuckets);
    if (nb < 1024) nb = 1024;
    nb = 1 << ceil_log2_l(nb);
    long bucket_bytes = 8L * (long)nb;
    if (inner_rel_bytes + (float)bucket_bytes > (float)hash_table_bytes) {
        long bucket_size = (long)tupsize + 8L;
        long lbuckets = 1L << ceil_log2_l(div_l(hash_table_bytes, bucket_size));
        if (max_pointers < lbuckets) lbuckets = max_pointers;
        nb = (int)lbuckets;
        nb = 1 << ceil_log2_l(nb);
        bucket_bytes = (long)nb * 8L;
        float dbatch = ceilf(inner_rel_bytes / (float)(hash_table_bytes - bucket_bytes));
        dbatch = fminf(dbatch, (float)max_pointers);
        int minbatch = f2i_x86(dbatch);
        int p2 = 1 << ceil_log2_l(minbatch);
        res.nbatches = p2 > 2 ? p2 : 2;
    }
    res.nbuckets = nb;
    return res;
}

/* Match mask of the ECs with members on both sides. */
template <int ECW = 0, class S>
GQ_HD int match_mask(const View<S> &q, const S &l, const S &r, uint64_t *m) {
    const int nw = ECW ? ECW : q.ecw;        /* ECW: class-mask words fixed at compile time (0: q.ecw) */
    uint64_t ml[GQ_MAX_ECW], mr[GQ_MAX_ECW];
    for (int w = 0; w < nw; w++) { ml[w] = 0; mr[w] = 0; }
    S t = l;
    while (!empty(t)) {
        int i = ctz(t);
        t = t - lowbit(t);
        for (int w = 0; w < nw; w++) ml[w] |= q.inc[i * nw + w];
    }
    t = r;
    while (!empty(t)) {
        int i = ctz(t);
        t = t - lowbit(t);
        for (int w = 0; w < nw; w++) mr[w] |= q.inc[i * nw + w];
    }
    int n = 0;
    for (int w = 0; w < nw; w++) { m[w] = ml[w] & mr[w]; n += popc(m[w]); }
    return n;
}

/* fkselc_of_ec (row_estimation.cuh:26-45): returns false when no FK applies. */
template <class S>
GQ_HD bool fk_sel(const View<S> &q, int e, const S &outer_m, const S &inner_m, float *out) {
    S ec = q.ec_relids[e];
    S t = outer_m;
    while (!empty(t)) {
        S o = lowbit(t);
        int idx = rank_below(ec, o);
        S match = inner_m & q.mem_fks[q.ec_moff[e] + idx];
        if (!empty(match)) { *out = q.rel_fksel[ctz(match)]; return true; }
        t = t - o;
    }
    return false;
}

/* estimate_ec_selectivity for a class with members on both sides
 * (row_estimation.cuh:48-86). */
template <class S>
GQ_HD float ec_selectivity(const View<S> &q, int e, const S &l, const S &r) {
    S ec = q.ec_relids[e];
    S ml = ec & l, mr = ec & r;
    float s;
    if (q.ec_hasfk[e]) {
        if (fk_sel(q, e, ml, mr, &s)) return s;
        if (fk_sel(q, e, mr, ml, &s)) return s;
    }
    int il = rank_below(ec, ml), ir = rank_below(ec, mr);
    return q.sels[q.ec_soff[e] + eq_class_index(il, ir, popc(ec))];
}

/* Visit the classes of mask `m` in ascending index. */
template <int ECW = 0, class S, class F>
GQ_HD void for_each_ec(const View<S> &q, const uint64_t *m, F f) {
    const int nw = ECW ? ECW : q.ecw;
    for (int w = 0; w < nw; w++) {
        uint64_t b = m[w];
        while (b) {
            int e = w * 64 + __builtin_ctzll(b);
            b &= b - 1;
            f(e);
        }
    }
}

/* estimate_join_selectivity (row_estimation.cuh:90-117) */
template <int ECW = 0, class S>
GQ_HD float join_selectivity(const View<S> &q, const S &l, const S &r, const uint64_t *m) {
    float sel = 1.0f;
    for_each_ec<ECW>(q, m, [&](int e) { sel *= ec_selectivity(q, e, l, r); });
    return sel;
}

/* estimate_join_rows (row_estimation.cuh:120-133) */
GQ_HD float join_rows(float sel, float lrows, float rrows) {
    float rows = sel * lrows * rrows;
    return rows > 1.0f ? roundf(rows) : 1.0f;
}

/* outer_match_frac and scan_frac (the inner_scan_frac of cost_nestloop and
 * cost_hashjoin, 2 / (match_count + 1)) are set when inner_unique only:
 * nothing else reads them */
struct Extra { bool inner_unique, indexed; float joinrows, outer_match_frac, match_count, scan_frac; };

/* ExecChooseHashTableSize depends on the inner relation alone (rows,
 * width) and the query's work_mem, so the memo keeps it with the relation:
 * bit 31 set, log2 nbuckets in bits 21-25, log2 nbatches in bits 26-30 and
 * the width in bits 0-20.  0 when the width needs more than 21 bits or a
 * size is not a power of two below 2^31 (then it is computed per join). */
GQ_HD uint32_t hash_pack(float rows, int width, int work_mem) {
    if (width < 0 || width >= (1 << 21)) return 0u;
    const HashSize hs = hash_table_size(rows, width, work_mem);
    if (hs.nbuckets <= 0 || (hs.nbuckets & (hs.nbuckets - 1)) || hs.nbatches <= 0 ||
        (hs.nbatches & (hs.nbatches - 1)))
        return 0u;
    return 0x80000000u | ((uint32_t)__builtin_ctz((unsigned)hs.nbatches) << 26) |
           ((uint32_t)__builtin_ctz((unsigned)hs.nbuckets) << 21) | (uint32_t)width;
}
GQ_HD HashSize hash_unpack(uint32_t hs) {
    HashSize r;
    r.nbuckets = 1 << ((hs >> 21) & 31u);
    r.nbatches = 1 << ((hs >> 26) & 31u);
    return r;
}

/* cost_nestloop (cost_postgres.cuh:110-275); qual_per_tuple is
 * cost_qual_eval's per-tuple cost (its startup is 0). */
template <class S>
GQ_HD void nestloop(const View<S> &q, const Rel &outer, float in_rows, float in_startup, float in_total,
                    const Extra &ex, float qual_per_tuple, float *o_startup, float *o_total) {
    float startup_cost = 0.0f, run_cost = 0.0f, ntuples;
    float outer_path_rows = outer.rows, inner_path_rows = in_rows;
    if (!q.p.enable_nestloop) startup_cost += q.p.disable_cost;
    startup_cost += in_startup + outer.startup;
    run_cost += outer.total - outer.startup;
    if (outer_path_rows > 1.0f) run_cost += (outer_path_rows - 1.0f) * in_startup;
    float inner_run_cost = in_total - in_startup;
    float inner_rescan_run_cost = inner_run_cost;
    if (ex.inner_unique) {
        float outer_matched_rows = rintf(outer_path_rows * ex.outer_match_frac);
        float outer_unmatched_rows = outer_path_rows - outer_matched_rows;
        const float inner_scan_frac = ex.scan_frac;
        ntuples = outer_matched_rows * inner_path_rows * inner_scan_frac;
        if (ex.indexed) {
            run_cost += inner_run_cost * inner_scan_frac;
            if (outer_matched_rows > 1.0f)
                run_cost += (outer_matched_rows - 1.0f) * inner_rescan_run_cost * inner_scan_frac;
            run_cost += div_pow2(outer_unmatched_rows * inner_rescan_run_cost, inner_path_rows,
                                 recip_pow2(inner_path_rows));
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
    startup_cost += 0.0f;   /* cost_qual_eval startup (always 0); keeps -0.0 -> +0.0 */
    float cpu_per_tuple = q.p.cpu_tuple_cost + qual_per_tuple;
    run_cost += cpu_per_tuple * ntuples;
    *o_startup = startup_cost;
    *o_total = startup_cost + run_cost;
}

/* __estimate_hash_bucketsize with the nbuckets-independent part precomputed
 * per member (cost_postgres.cuh:701-775): fb = (float)nbuckets and
 * inv_fb = 1.0f / fb come from the join, 1.0f / ndistinct and the skew
 * quotient mcvfreq / avgfreq from the member (the quotients the reference
 * computes, rounded once; a member without skew multiplies by 1.0f, which
 * changes no bit). */
template <class S>
GQ_HD float bucketsize(const View<S> &q, int mi, float fb, float inv_fb) {
    if (q.mem_flags[mi] & GQ_MF_BUCKET_DEFAULT) return 0.1f;
    float est = q.mem_nd[mi] > fb ? inv_fb : q.mem_inv_nd[mi];
    est *= q.mem_skew[mi];
    if (est < 1.0e-6f) est = 1.0e-6f;
    else if (est > 1.0f) est = 1.0f;
    return est;
}

/* estimate_hash_bucket_stats (cost_postgres.cuh:620-668) folds one inner
 * member: the minima do not depend on the order the members come in. */
template <class S>
GQ_HD void bucket_fold(const View<S> &q, int mi, float fb, float inv_fb, float *bsize, float *mcvfreq) {
    const float t = bucketsize(q, mi, fb, inv_fb);
    if (*bsize > t) *bsize = t;
    if (*mcvfreq > q.mem_mcv[mi]) *mcvfreq = q.mem_mcv[mi];
}

/* cost_hashjoin (cost_postgres.cuh:287-470) with the inner's hash table size
 * `hs` (virtualbuckets = nbuckets * nbatches) and the bucket statistics
 * (bsize, mcvfreq) of its crossing members already computed (join_prov). */
template <class S>
GQ_HD void hashjoin(const View<S> &q, const Rel &outer, const Rel &inner, const Extra &ex, int n_quals,
                    float qual_per_tuple, HashSize hs, float virtualbuckets, float inv_vb, float bsize,
                    float mcvfreq, float *o_startup, float *o_total) {
    float startup_cost = 0.0f, run_cost = 0.0f;
    float outer_path_rows = outer.rows, inner_path_rows = inner.rows;
    if (!q.p.enable_hashjoin) startup_cost += q.p.disable_cost;
    startup_cost += outer.startup;
    run_cost += outer.total - outer.startup;
    startup_cost += inner.total;
    startup_cost += (q.p.cpu_operator_cost * (float)n_quals + q.p.cpu_tuple_cost) * inner_path_rows;
    run_cost += q.p.cpu_operator_cost * (float)n_quals * outer_path_rows;
    if (hs.nbatches > 1) {
        float outerpages = page_size(outer_path_rows, outer.width);
        float innerpages = page_size(inner_path_rows, inner.width);
        startup_cost += q.p.seq_page_cost * innerpages;
        run_cost += q.p.seq_page_cost * (innerpages + 2.0f * outerpages);
    }
    if (relation_byte_size(clamp_row_est(inner_path_rows * mcvfreq), inner.width) >
        (float)((long)q.p.work_mem * 1024L))
        startup_cost += q.p.disable_cost;
    float hashjointuples;
    startup_cost += 0.0f;   /* hash_qual_cost.startup (always 0), both branches */
    if (ex.inner_unique) {
        float outer_matched_rows = rintf(outer_path_rows * ex.outer_match_frac);
        const float inner_scan_frac = ex.scan_frac;
        run_cost += qual_per_tuple * outer_matched_rows *
                    clamp_row_est(inner_path_rows * bsize * inner_scan_frac) * 0.5f;
        run_cost += qual_per_tuple * (outer_path_rows - outer_matched_rows) *
                    clamp_row_est(div_pow2(inner_path_rows, virtualbuckets, inv_vb)) * 0.05f;
        hashjointuples = outer_matched_rows;
    } else {
        run_cost += qual_per_tuple * outer_path_rows * clamp_row_est(inner_path_rows * bsize) * 0.5f;
        hashjointuples = ex.joinrows;
    }
    run_cost += q.p.cpu_tuple_cost * hashjointuples;
    *o_startup = startup_cost;
    *o_total = startup_cost + run_cost;
}

/* One join: rows (estimate_join_rows) and cost (calc_join_cost,
 * gpuqo_cost.cuh:76-158) of outer `l` joined with inner `r`. */
/* The crossing classes of one join, as join_prov sees them.
 *
 * MaskProv: the classes of mask `m` (ascending index); the selectivity is
 * estimate_join_selectivity over them and the inner members are found by
 * rank within each class.
 *
 * EdgeProv: a split of a tree query at edge (p, c): its crossing classes
 * are those on that edge, so the selectivity and the inner member slots
 * depend only on (c, orientation) and are precomputed on the host with this
 * same code (ec_selectivity on the one-member sides, the product taken in
 * ascending class order). */
template <int ECW>
struct MaskProv {
    const uint64_t *m;
    int nq;
    template <class S>
    GQ_HD float selectivity(const View<S> &q, const S &l, const S &r) const { return join_selectivity<ECW>(q, l, r, m); }
    GQ_HD int n_quals() const { return nq; }
    /* One pass over the crossing classes: the selectivity (product in
     * ascending class order) and, with `buckets`, the bucket statistics of
     * the inner members. */
    template <class S>
    GQ_HD float scan(const View<S> &q, const S &l, const S &r, bool buckets, float fb, float inv_fb, float *bsize,
                     float *mcvfreq) const {
        float sel = 1.0f;
        if (q.pairs) {
            for_each_ec<ECW>(q, m, [&](int e) {
                const PairRec *pr = q.pair_rec + 3 * e;
                const PairRec c = pr[0], m0 = pr[1], m1 = pr[2];   /* three independent reads */
                const bool lo_outer = bit(l, (int)c.z);
                sel *= lo_outer ? c.x : c.y;
                if (buckets) {
                    const PairRec mm = lo_outer ? m1 : m0;           /* the inner member */
                    float est = mm.x > fb ? inv_fb : mm.y;
                    est *= mm.z;
                    if (est < 1.0e-6f) est = 1.0e-6f;
                    else if (est > 1.0f) est = 1.0f;
                    if (*bsize > est) *bsize = est;
                    if (*mcvfreq > mm.w) *mcvfreq = mm.w;
                }
            });
        } else {
            for_each_ec<ECW>(q, m, [&](int e) {
                sel *= ec_selectivity(q, e, l, r);
                if (buckets) {
                    S ec = q.ec_relids[e];
                    S mr = ec & r;
                    while (!empty(mr)) {
                        S in_id = lowbit(mr);
                        bucket_fold(q, q.ec_moff[e] + rank_below(ec, in_id), fb, inv_fb, bsize, mcvfreq);
                        mr = mr - in_id;
                    }
                }
            });
        }
        return sel;
    }
    /* f(member slot) for each member of `inner` in each class, in class order */
    template <class S, class F>
    GQ_HD void for_each_inner(const View<S> &q, const S &inner, F f) const {
        for_each_ec<ECW>(q, m, [&](int e) {
            S ec = q.ec_relids[e];
            S mr = ec & inner;
            while (!empty(mr)) {
                S in_id = lowbit(mr);
                f(q.ec_moff[e] + rank_below(ec, in_id));
                mr = mr - in_id;
            }
        });
    }
};

struct EdgeProv {
    float sel;
    int nq, n_mi;
    const uint16_t *mi;
    template <class S>
    GQ_HD float selectivity(const View<S> &, const S &, const S &) const { return sel; }
    template <class S>
    GQ_HD float scan(const View<S> &q, const S &, const S &, bool buckets, float fb, float inv_fb, float *bsize,
                     float *mcvfreq) const {
        if (buckets) {
            if (q.pairs) {
                /* member m of a two-member class e is m = 2e + j, and its
                 * bucket constants are pair_rec[3e + 1 + j] (one 16-byte
                 * read; the 0.1 default members give exactly 0.1) */
                for (int i = 0; i < n_mi; i++) {
                    const int m = (int)mi[i];
                    const PairRec mm = q.pair_rec[3 * (m >> 1) + 1 + (m & 1)];
                    float est = mm.x > fb ? inv_fb : mm.y;
                    est *= mm.z;
                    if (est < 1.0e-6f) est = 1.0e-6f;
                    else if (est > 1.0f) est = 1.0f;
                    if (*bsize > est) *bsize = est;
                    if (*mcvfreq > mm.w) *mcvfreq = mm.w;
                }
            } else {
                for (int i = 0; i < n_mi; i++) bucket_fold(q, (int)mi[i], fb, inv_fb, bsize, mcvfreq);
            }
        }
        return sel;
    }
    GQ_HD int n_quals() const { return nq; }
    template <class S, class F>
    GQ_HD void for_each_inner(const View<S> &, const S &, F f) const {
        for (int i = 0; i < n_mi; i++) f((int)mi[i]);
    }
};

/* calc_join_cost (gpuqo_cost.cuh:76-158) under the simple and C_out cost
 * functions, whose startup costs are 0: the nested loop costs
 * outer.total + outer.rows * inner.total (cost_simple.cuh:44-56) or +inf
 * (cost_cout.cuh:40-52); the hash join, when enabled, joinrows +
 * inner.total + outer.total (cost_simple.cuh:58-70 with HASHJOIN_COEFF
 * 1.0f, an exact product; cost_cout.cuh:54-66); the index nested loop, for
 * a non-composite single inner relation, the nested loop over each crossing
 * member's index (its cost's total), in class order.  Each later candidate
 * replaces the current one when strictly cheaper. */
template <class S, class Prov>
GQ_HD Rel join_plain(const View<S> &q, const S &l, const Rel &L, const S &r, const Rel &R, const Prov &pv) {
    Rel out;
    float bsize = 1.0f, mcvfreq = 1.0f;
    out.rows = join_rows(pv.scan(q, l, r, false, 0.0f, 0.0f, &bsize, &mcvfreq), L.rows, R.rows);
    out.width = L.width + R.width;
    out.hs = 0u;
    const bool simple = q.p.cost_fn == COST_FN_SIMPLE;
    float bt = simple ? L.total + L.rows * R.total : __builtin_inff();
    if (q.p.enable_hashjoin) {
        const float ht = out.rows + R.total + L.total;
        if (ht < bt) bt = ht;
    }
    if (simple && popc(r) == 1 && !bit(q.composite, ctz(r))) {
        pv.for_each_inner(q, r, [&](int mi) {
            if (q.mem_flags[mi] & GQ_MF_INDEX) {
                const float it = L.total + L.rows * q.mem_idx_total[mi];
                if (it < bt) bt = it;
            }
        });
    }
    out.startup = 0.0f;
    out.total = bt;
    return out;
}

/* One join: rows (estimate_join_rows) and cost (calc_join_cost,
 * gpuqo_cost.cuh:76-158) of outer `l` joined with inner `r`. */
template <class S, class Prov>
GQ_HD Rel join_prov(const View<S> &q, const S &l, const Rel &L, const S &r, const Rel &R, const Prov &pv) {
    if (__builtin_expect(q.p.cost_fn != COST_FN_POSTGRES, 0)) return join_plain(q, l, L, r, R, pv);
    Rel out;
    /* the inner's hash table size first (it depends on the inner alone), so
     * one pass over the crossing classes gives the selectivity and the
     * bucket statistics together */
    const bool hj = q.p.enable_hashjoin != 0;
    HashSize hsz;
    hsz.nbuckets = 0;
    hsz.nbatches = 0;
    float vbuckets = 0.0f, inv_vb = 0.0f, fb = 0.0f, inv_fb = 0.0f, bsize = 1.0f, mcvfreq = 1.0f;
    if (hj) {
        hsz = R.hs ? hash_unpack(R.hs) : hash_table_size(R.rows, R.width, q.p.work_mem);
        vbuckets = (float)hsz.nbuckets * (float)hsz.nbatches;
        inv_vb = recip_pow2(vbuckets);         /* sizes are powers of two: exact reciprocals */
        fb = (float)f2i_x86(vbuckets);         /* the int nbuckets of estimate_hash_bucket_stats */
        inv_fb = recip_pow2(fb);
        if (inv_fb == 0.0f) inv_fb = 1.0f / fb;
    }
    out.rows = join_rows(pv.scan(q, l, r, hj, fb, inv_fb, &bsize, &mcvfreq), L.rows, R.rows);
    out.width = L.width + R.width;
    out.hs = 0u;
    const int n_quals = pv.n_quals();
    float qual_pt = q.qualtab[n_quals];
    bool single = popc(r) == 1;
    Extra ex;
    ex.inner_unique = false;
    if (single) {                                   /* is_inner_unique :43-74 */
        pv.for_each_inner(q, r, [&](int mi) {
            if ((q.mem_flags[mi] & (GQ_MF_INDEX | GQ_MF_UNIQUE)) == (GQ_MF_INDEX | GQ_MF_UNIQUE))
                ex.inner_unique = true;
        });
    }
    ex.indexed = false;
    ex.joinrows = out.rows;
    ex.match_count = R.rows;
    ex.outer_match_frac = 0.0f;
    ex.scan_frac = 0.0f;
    if (ex.inner_unique) {
        ex.outer_match_frac = out.rows / (L.rows * R.rows);
        ex.scan_frac = 2.0f / (ex.match_count + 1.0f);
    }
    float bs, bt;
    nestloop(q, L, R.rows, R.startup, R.total, ex, qual_pt, &bs, &bt);
    if (hj) {
        float hs, ht;
        hashjoin(q, L, R, ex, n_quals, qual_pt, hsz, vbuckets, inv_vb, bsize, mcvfreq, &hs, &ht);
        if (ht < bt) { bs = hs; bt = ht; }
    }
    if (single && !bit(q.composite, ctz(r))) {       /* index nested loop :116-147 */
        ex.indexed = true;
        pv.for_each_inner(q, r, [&](int mi) {
            if (q.mem_flags[mi] & GQ_MF_INDEX) {
                float is, it;
                nestloop(q, L, q.mem_idx_rows[mi], q.mem_idx_startup[mi], q.mem_idx_total[mi], ex, qual_pt,
                         &is, &it);
                if (it < bt) { bs = is; bt = it; }
            }
        });
    }
    out.startup = bs;
    out.total = bt;
    return out;
}

/* The same join with the crossing-class mask `m` (n_quals classes) given. */
template <int ECW = 0, class S>
GQ_HD Rel join_mask(const View<S> &q, const S &l, const Rel &L, const S &r, const Rel &R, const uint64_t *m,
                    int n_quals) {
    MaskProv<ECW> pv;
    pv.m = m;
    pv.nq = n_quals;
    return join_prov(q, l, L, r, R, pv);
}

template <int ECW = 0, class S>
GQ_HD Rel join(const View<S> &q, const S &l, const Rel &L, const S &r, const Rel &R) {
    uint64_t m[GQ_MAX_ECW];
    int n_quals = match_mask<ECW>(q, l, r, m);
    return join_mask<ECW>(q, l, L, r, R, m, n_quals);
}

}  // namespace gq

#endif /* GQ_COST_MODEL_H */
