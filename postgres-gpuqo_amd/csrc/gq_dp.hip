/*
 * gq_dp.hip -- the device side of the join-order search on CDNA4 (gfx950,
 * wave64): DPsub / MPDP and DPsize level loops, their batched and
 * step-engine forms, and the plan walk.
 *
 * One DP covers n <= 32 relations (relation i = bit i of a uint32_t set).
 * Its level lists come from one of two producers:
 *   k_filter_all   every integer in [0, 2^n) is a candidate; a wave tests 64
 *                  consecutive ones for connectivity (O(1) for trees, a
 *                  closure from the lowest member otherwise), sets their
 *                  bits in the connectivity bitmap and compacts the
 *                  connected ones of each size through LDS into the level
 *                  lists, one global position reservation per workgroup and
 *                  level (lists are increasing within a workgroup's chunk;
 *                  chunks land in completion order -- no result depends on
 *                  the order of a level list);
 *   expansion      tree graphs: each connected set of k+1 relations is built
 *                  once from the set of its k lowest-numbered members
 *                  (k_tree_start, the expansion workgroups of k_eval_tree,
 *                  k_expand), so no 2^n pass runs.
 * Each level is then one launch (no host synchronisation between levels):
 *   k_eval_tree       trees: 2(k-1) splits per set, per-edge cost tables;
 *   k_eval_wg         general graphs: MPDP block decomposition or DPsub
 *                     subset scan, valid pairs queued in LDS;
 *   k_big_scan/eval   MPDP sets with very large biconnected blocks;
 *   k_eval_dpsub_chunk + k_finalize_dpsub   DPsub levels with > 4096
 *                     candidate splits per set;
 *   k_ds_*            DPsize: tile-parallel pair scan in the CPU's
 *                     creation order.
 * k_tree walks the memo from the root to the plan.
 *
 * Memos (chosen per DP): dense -- entry S at index S, 16 B of value plus a
 * 4 B left child, up to 2^32 entries in HBM; rank -- tree graphs, one slot
 * per connected set computed from the set; open addressing -- tree graphs,
 * linear probing.  Ties and NaN costs follow the reference CPU oracles: per
 * set the first valid split in their enumeration order is kept unless a
 * later one is strictly cheaper (gpuqo_cpu_common.cuh:189-211), which the
 * kernels reproduce with an associative (first valid, first minimum)
 * reduction over packed (cost bits, rank) keys.
 *
 * Reference counterparts: gpuqo_dpsub_common.cu:177-392 (level loop),
 * gpuqo_dpsub_filtered_kernels.cuh:30-240 (unrank/filter/evaluate),
 * gpuqo_dpsub_enum_all_subs.cuh:21-72 (all-subsets splits),
 * gpuqo_dpsub_bicc.cuh:35-313 (MPDP), gpuqo_cpu_dpsub_bicc.cu:65-236 (the
 * block order reproduced here), gpuqo_dpsize.cu:292-671 and
 * gpuqo_cpu_dpsize.cu:27-93 (DPsize and its creation order).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gq_device.h"
#include "gq_dp.h"
#include "gq_internal.h"
#include "gq_wide.h"

namespace gq {

#define WAVE 64
#define BLOCK 256
#define XCDS 8u     /* XCDs of an MI355X; workgroups are dispatched to them round robin */
#define BMAX 128                 /* sets per workgroup batch            */
#define FILT_CHUNK 4096          /* candidates per filter block pass    */
#define QCAP 1024                /* DPsub valid-pair queue per round    */
/* Dynamic LDS holds 64-bit slots updated with LDS atomics, which need
 * natural alignment; kernels using it declare no static __shared__. */
#define GQ_LDS_ALIGN __attribute__((aligned(16)))


struct DevBufs {
    float4 *vals;          /* memo: rows, startup, total, width(bits) */
    uint32_t *left;        /* memo: left child set                    */
    uint32_t *conn;        /* connectivity bitmap, 2^n bits           */
    uint32_t *lvl_sets;    /* sets of the current level               */
    unsigned long long *lvl_best;
    uint32_t *lvl_first;
    uint32_t *counters;    /* [0..32] sets per level, [40] pairs lo, [41] pairs hi */
    uint32_t *blob;
    DPNode *nodes;
};

/* ---------------------------------------------------------------- helpers */

__device__ __forceinline__ View<uint32_t> make_view(const BlobHdr &h, const uint32_t *s) {
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

__device__ __forceinline__ void stage(const uint32_t *g, uint32_t *s, int words) {
    for (int i = threadIdx.x; i < words; i += blockDim.x) s[i] = g[i];
    __syncthreads();
}

__device__ __forceinline__ bool is_conn(const uint32_t *conn, uint32_t x) {
    return (conn[x >> 5] >> (x & 31)) & 1u;
}

/* Memo records (rel_of / record_of, gq_device.h) of slot x. */
__device__ __forceinline__ Rel load_rel(const float4 *vals, uint32_t x) { return rel_of(vals[x]); }

/* pack = false: for latency-bound levels, where the winning lane alone would
 * run hash_table_size on the critical path. */
__device__ __forceinline__ void store_rel(float4 *vals, uint32_t x, const Rel &r, int work_mem, bool pack = true) {
    vals[x] = record_of(r, work_mem, pack);
}

/* base relation i of the blob into memo slot x */
__device__ __forceinline__ void store_base(float4 *vals, uint32_t x, const BlobHdr &h, const uint32_t *blob, int i) {
    const float *b = (const float *)(blob + h.o_base) + 4 * i;
    Rel r;
    r.rows = b[0];
    r.startup = b[1];
    r.total = b[2];
    r.width = __float_as_int(b[3]);
    store_rel(vals, x, r, h.p.work_mem);
}

/* The memo of one DP.  Dense (keys == nullptr, rank == nullptr): the slot
 * of set S is S.  Open addressing (tree graphs): keys[slot] holds S (0 =
 * empty), linear probing from hash(S); every connected set is inserted
 * before its values are read, so a probe ends at its key.  A probe that
 * meets an empty slot or runs over the whole table sets bit 2 of *err and
 * returns slot 0.  Rank (tree graphs, rank != nullptr): the slot of
 * connected set S is rank_slot(rank, rn, S), one slot per connected set. */
struct MemoRef {
    float4 *vals;
    uint32_t *left;
    const uint32_t *keys;
    uint32_t mask;
    uint32_t *err;
    const uint32_t *rank;
    int rn;
};

/* Rank memo of a tree graph whose every vertex has a lower-numbered parent
 * (BlobHdr::is_tree): connected set S with lowest member t has slot
 *     off[t] + sum over u in S - t of Q[t][u]
 * (rk = off[rn], Q[rn][rn], inv[rn], tz[rn], then the byte tables of W; the
 * tables are built by attach_rank_tables, which shows that this numbers the
 * connected sets 0, 1, ... without gaps).  The sum over S - t is
 * W(S) / w[t] - 1 with W(S) = sum over u in S of Q[0][u] (w[t] = Q[0][t],
 * w[0] = 1), four table lookups instead of a loop over S. */
__device__ __forceinline__ uint32_t rank_wsum(const uint32_t *rk, int rn, uint32_t x) {
    const uint32_t *bt = rk + rn + rn * rn + 2 * rn;
    uint32_t w = bt[x & 255u];
    if (rn > 8) w += bt[256 + ((x >> 8) & 255u)];
    if (rn > 16) w += bt[512 + ((x >> 16) & 255u)];
    if (rn > 24) w += bt[768 + (x >> 24)];
    return w;
}
/* 1 + the sum over u in x - t of Q[t][u], t the lowest member of the
 * connected set x */
__device__ __forceinline__ uint32_t rank_digits(const uint32_t *rk, int rn, uint32_t x, int t) {
    const uint32_t *inv = rk + rn + rn * rn, *tz = inv + rn;
    return (rank_wsum(rk, rn, x) >> tz[t]) * inv[t];
}
__device__ __forceinline__ uint32_t rank_slot(const uint32_t *rk, int rn, uint32_t S) {
    const int t = __builtin_ctz(S);
    return rk[t] + rank_digits(rk, rn, S, t) - 1u;
}

__device__ __forceinline__ uint32_t memo_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t memo_slot(const MemoRef &m, uint32_t S) {
    if (m.rank) return rank_slot(m.rank, m.rn, S);
    if (!m.keys) return S;
    uint32_t h = memo_hash(S) & m.mask;
    for (uint32_t i = 0; i <= m.mask; i++) {
        const uint32_t key = m.keys[h];
        if (key == S) return h;
        if (key == 0u) break;
        h = (h + 1u) & m.mask;
    }
    atomicOr(m.err, 4u);
    return 0u;
}

/* Inserts S (absent or present); returns its slot and whether this call
 * inserted it, or ~0u when the table is full. */
__device__ __forceinline__ uint32_t memo_insert(uint32_t *keys, uint32_t mask, uint32_t S, bool *fresh) {
    uint32_t h = memo_hash(S) & mask;
    for (uint32_t i = 0; i <= mask; i++) {
        const uint32_t old = atomicCAS(&keys[h], 0u, S);
        if (old == 0u) {
            *fresh = true;
            return h;
        }
        if (old == S) {
            *fresh = false;
            return h;
        }
        h = (h + 1u) & mask;
    }
    *fresh = false;
    return ~0u;
}

/* closure of `from` inside `subset` over the adjacency table (gpuqo_filter.cuh:113-146) */
__device__ __forceinline__ uint32_t grow(uint32_t from, uint32_t subset, const uint32_t *adj) {
    uint32_t V = 0, N = from;
    while (N) {
        int i = __builtin_ctz(N);
        V |= 1u << i;
        N |= adj[i];
        N &= subset;
        N &= ~V;
    }
    return V;
}

/* index of the member of x that has n members below it (n < popcount(x)):
 * a binary search on popcounts of halves, 5 steps whatever the size of x */
__device__ __forceinline__ int select_bit(uint32_t x, uint32_t n) {
    int p = 0;
    uint32_t c = __builtin_popcount(x & 0xffffu);
    if (n >= c) { n -= c; x >>= 16; p += 16; }
    c = __builtin_popcount(x & 0xffu);
    if (n >= c) { n -= c; x >>= 8; p += 8; }
    c = __builtin_popcount(x & 0xfu);
    if (n >= c) { n -= c; x >>= 4; p += 4; }
    c = __builtin_popcount(x & 0x3u);
    if (n >= c) { n -= c; x >>= 2; p += 2; }
    return (n >= (x & 1u)) ? p + 1 : p;
}

__device__ __forceinline__ uint32_t pdep32(uint32_t j, uint32_t mask) {
    uint32_t res = 0;
    while (mask) {
        uint32_t b = mask & (0u - mask);
        if (j & 1u) res |= b;
        j >>= 1;
        mask &= mask - 1;
    }
    return res;
}

/* C(n, k) at [n * 33 + k], n, k <= 32 (uploaded by ctx_init) */
__constant__ uint32_t c_binom[33 * 33];

/* The t-th proper non-empty subset of k positions (t < 2^k - 2) in order of
 * size, then of numeric value within a size (combinadic unranking).  The
 * DPsub kernels enumerate a k-set's splits in this order -- each split keeps
 * its numeric rank, so the first-valid and first-minimum keys are those of
 * the reference's increasing-rank order -- because neighbouring lanes then
 * cost splits with the same left size, i.e. (on dense graphs) the same
 * number of crossing classes, and do not wait on each other. */
__device__ __forceinline__ uint32_t split_by_size(uint32_t t, int k) {
    const uint32_t *row = c_binom + k * 33;
    int a = 1;
    uint32_t c = row[1];
    while (t >= c && a < k - 1) {
        t -= c;
        c = row[++a];
    }
    uint32_t m = 0;
    for (int pos = k - 1; pos >= 0 && a > 0; pos--) {
        const uint32_t b = c_binom[pos * 33 + a];
        if (t >= b) {
            m |= 1u << pos;
            t -= b;
            a--;
        }
    }
    return m;
}

/* Adds the wave's pair count to counters[40..41]: one atomic per wave.  All
 * lanes of the wave must call it. */
__device__ __forceinline__ void add_pairs(uint32_t *counters, uint32_t n) {
    for (int off = WAVE / 2; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((threadIdx.x & (WAVE - 1)) == 0 && n) atomicAdd((unsigned long long *)(counters + 40), (unsigned long long)n);
}

/* ---------------------------------------------------------------- kernels */

__global__ void k_clear(uint32_t *conn, uint32_t words, uint32_t *counters) {
    uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = tid; i < words; i += gridDim.x * blockDim.x) conn[i] = 0;
    if (tid < 64) counters[tid] = 0;
}

__global__ void k_base(BlobHdr h, const uint32_t *blob, float4 *vals, uint32_t *left, uint32_t *conn) {
    int i = threadIdx.x;
    if (i < h.n_rels) {
        uint32_t s = 1u << i;
        store_base(vals, s, h, blob, i);
        left[s] = 0;
        atomicOr(&conn[s >> 5], 1u << (s & 31));
    }
}

/* All levels at once: every integer s in [0, 2^n) is a candidate set; the
 * connected ones (singletons included) are marked in the bitmap -- each wave
 * covers 64 consecutive integers, i.e. two whole bitmap words, so the words
 * are written without atomics -- and sets of two or more relations are
 * appended to the list of their level (offset lvl_off[k]) with one atomic
 * per (wave, level). */
/* One chunk of FILT_CHUNK candidate sets (integers ch*FILT_CHUNK + i):
 * connected ones are marked in the bitmap -- a wave owns 64 consecutive
 * candidates, i.e. two whole bitmap words, written without atomics -- and
 * those of two or more members are compacted in LDS, then appended to their
 * level list with one global atomic per (chunk, level).  Trees test
 * connectivity in O(1): S is connected iff |S & children(S)| = |S| - 1,
 * children(S) coming from a per-byte table `lut`; other graphs grow the
 * closure of the lowest member. */
struct FilterSmem {
    uint32_t adj[32];
    uint32_t lut[4 * 256];
    uint32_t kept[FILT_CHUNK];
    uint32_t lcnt[33], lbase[33], nkept;
};

__device__ void filter_setup(FilterSmem &F, const BlobHdr &h, const uint32_t *blob) {
    const int tid = threadIdx.x, n = h.n_rels;
    if (tid < 32) F.adj[tid] = tid < n ? blob[h.o_adj + tid] : 0u;
    __syncthreads();
    if (h.is_tree)
        for (int e = tid; e < 4 * 256; e += BLOCK) {
            int byte = e >> 8, x = e & 255;
            uint32_t m = 0;
            for (int i = 0; i < 8; i++) {
                int v = byte * 8 + i;
                if (((x >> i) & 1) && v < n) m |= F.adj[v] & ~((2u << v) - 1u);
            }
            F.lut[e] = m;
        }
    __syncthreads();
}

__device__ void filter_chunk(FilterSmem &F, const BlobHdr &h, uint64_t ch, const uint32_t *lvl_off,
                             uint32_t *lvl_sets, unsigned long long *lvl_best, uint32_t *lvl_first, uint32_t *conn,
                             uint32_t conn_words, uint32_t *counters) {
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const bool tree = h.is_tree != 0;
    const uint64_t total = 1ull << h.n_rels;
    if (tid < 33) F.lcnt[tid] = 0;
    if (tid == 0) F.nkept = 0;
    __syncthreads();
    for (uint32_t o = (uint32_t)(tid / WAVE) * WAVE; o < FILT_CHUNK; o += BLOCK) {
        const uint64_t s64 = ch * FILT_CHUNK + o + lane;
        const uint32_t s = (uint32_t)s64;
        bool keep = false;
        if (s64 < total && s != 0u) {
            if (tree) {
                uint32_t chs = F.lut[s & 255u] | F.lut[256 + ((s >> 8) & 255u)] | F.lut[512 + ((s >> 16) & 255u)] |
                               F.lut[768 + (s >> 24)];
                keep = __builtin_popcount(s & chs) == __builtin_popcount(s) - 1;
            } else {
                keep = grow(s & (0u - s), s, F.adj) == s;
            }
        }
        unsigned long long m = __ballot(keep);
        const uint64_t wd = (ch * FILT_CHUNK + o) / 32;
        if (lane == 0 && wd < conn_words) conn[wd] = (uint32_t)m;
        if (lane == 0 && wd + 1 < conn_words) conn[wd + 1] = (uint32_t)(m >> 32);
        const bool pend = keep && __builtin_popcount(s) >= 2;
        unsigned long long pm = __ballot(pend);
        if (pm) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&F.nkept, (uint32_t)__popcll(pm));
            base = __shfl(base, 0);
            if (pend) {
                F.kept[base + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull))] = s;
                atomicAdd(&F.lcnt[__builtin_popcount(s)], 1u);
            }
        }
    }
    __syncthreads();
    if (tid < 33) {
        F.lbase[tid] = F.lcnt[tid] ? atomicAdd(&counters[tid], F.lcnt[tid]) : 0u;
        F.lcnt[tid] = 0;
    }
    __syncthreads();
    const uint32_t nk = F.nkept;
    for (uint32_t i = tid; i < nk; i += BLOCK) {
        const uint32_t s = F.kept[i];
        const int k = __builtin_popcount(s);
        const uint32_t pos = lvl_off[k] + F.lbase[k] + atomicAdd(&F.lcnt[k], 1u);
        lvl_sets[pos] = s;
        lvl_best[pos] = ~0ull;
        lvl_first[pos] = ~0u;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(BLOCK) k_filter_all(BlobHdr h, const uint32_t *blob, const uint32_t *lvl_off,
                                                       uint32_t *lvl_sets, unsigned long long *lvl_best,
                                                       uint32_t *lvl_first, uint32_t *conn, uint32_t conn_words,
                                                       uint32_t *counters) {
    __shared__ FilterSmem F;
    filter_setup(F, h, blob);
    const uint64_t nchunks = ((1ull << h.n_rels) + FILT_CHUNK - 1) / FILT_CHUNK;
    for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x)
        filter_chunk(F, h, ch, lvl_off, lvl_sets, lvl_best, lvl_first, conn, conn_words, counters);
}

/* DPsub, split-parallel: a wave owns `ipl * 64` consecutive split ranks of
 * one set; lane t scans ranks [j0 + t*ipl, j0 + (t+1)*ipl) in order.  Each
 * lane keeps (first valid rank, NaN flag) and the first minimum among
 * non-NaN costs; both reduce associatively across lanes and waves.  Five
 * waves per SIMD (at most 96 VGPRs, no spills; left alone the compiler
 * takes 103 and four waves: clique-18 DPsub 29.2 against 31.1 ms). */
__global__ void __launch_bounds__(BLOCK, 5) k_eval_dpsub_chunk(BlobHdr h, const uint32_t *blob, int k, uint32_t cps,
                                                             uint32_t ipl, const uint32_t *lvl_sets,
                                                             unsigned long long *lvl_best, uint32_t *lvl_first,
                                                             const float4 *vals, const uint32_t *conn,
                                                             uint32_t *counters) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    uint64_t nwaves = (uint64_t)counters[k] * cps;
    if ((uint64_t)blockIdx.x * (BLOCK / WAVE) >= nwaves) return;
    stage(blob, smem, h.words);
    View<uint32_t> q = make_view(h, smem);
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t T = (1u << k) - 2u;
    uint64_t wstride = (uint64_t)gridDim.x * (BLOCK / WAVE);
    uint32_t npairs = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE; w < nwaves; w += wstride) {
        uint32_t idx = (uint32_t)(w / cps), part = (uint32_t)(w % cps);
        uint32_t S = lvl_sets[idx];
        /* the wave's ipl * WAVE consecutive splits in size order, lane-strided */
        const uint64_t t0 = (uint64_t)part * ipl * WAVE + (uint64_t)lane;
        uint32_t first = ~0u;
        unsigned long long best = ~0ull;
        for (uint32_t s = 0; s < ipl; s++) {
            const uint64_t t = t0 + (uint64_t)s * WAVE;
            if (t >= T) break;
            const uint32_t j = split_by_size((uint32_t)t, k);
            const uint32_t l = pdep32(j, S), r = S ^ l;
            if (is_conn(conn, l) && is_conn(conn, r)) {
                Rel L = load_rel(vals, l), R = load_rel(vals, r);
                Rel jr = join(q, l, L, r, R);
                npairs++;
                bool nan = jr.total != jr.total;
                const uint32_t f = (j << 1) | (nan ? 1u : 0u);
                if (f < first) first = f;
                if (!nan) {
                    unsigned long long key = ((unsigned long long)ord_key(jr.total) << 32) | j;
                    if (key < best) best = key;
                }
            }
        }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            uint32_t f2 = __shfl_xor(first, off);
            unsigned long long b2 = __shfl_xor(best, off);
            first = f2 < first ? f2 : first;
            best = b2 < best ? b2 : best;
        }
        if (lane == 0) {
            if (first != ~0u) atomicMin(&lvl_first[idx], first);
            if (best != ~0ull) atomicMin(&lvl_best[idx], best);
        }
    }
    add_pairs(counters, npairs);
}

/* Winner of each set after k_eval_dpsub_chunk: the first valid split if its
 * cost is NaN, else the first minimum.  The join is recomputed. */
__global__ void __launch_bounds__(BLOCK) k_finalize_dpsub(BlobHdr h, const uint32_t *blob, int k,
                                                           const uint32_t *lvl_sets,
                                                           const unsigned long long *lvl_best,
                                                           const uint32_t *lvl_first, float4 *vals,
                                                           uint32_t *left, const uint32_t *counters) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    uint32_t count = counters[k];
    if ((uint64_t)blockIdx.x * blockDim.x >= count) return;
    stage(blob, smem, h.words);
    View<uint32_t> q = make_view(h, smem);
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < count; idx += gridDim.x * blockDim.x) {
        uint32_t S = lvl_sets[idx];
        uint32_t f = lvl_first[idx];
        if (f == ~0u) continue;
        uint32_t rank = (f & 1u) ? (f >> 1) : (uint32_t)(lvl_best[idx] & 0xffffffffull);
        uint32_t l = pdep32(rank, S), r = S ^ l;
        Rel j = join(q, l, load_rel(vals, l), r, load_rel(vals, r));
        store_rel(vals, S, j, q.p.work_mem);
        left[S] = l;
    }
}


/* Blocks of connected set S in the emission order of find_blocks
 * (gpuqo_cpu_dpsub_bicc.cu:65-169): DFS from the lowest vertex, neighbours
 * ascending; a block is emitted when a tree edge (p,c) finishes with
 * low[c] >= disc[p]; its vertices are p plus the vertices discovered at or
 * after c still on the vertex stack, i.e. c and the vertices above it on
 * the stack.  `dfs` is this set's column of a [32][stride] LDS array
 * (disc | low << 8 | parent << 16); `own` and `subb` are BSET_W-word rows.
 * Besides blocks[0..nblk) it fills, for block_side:
 *   own[v]   v not the DFS root: v and the vertices hanging off v outside
 *            the block where v is not the root (the blocks rooted at v and,
 *            recursively, what hangs off those); own[root] = S;
 *   subb[b]  the union of own[v] over the members of block b other than its
 *            root, i.e. the vertices below block b -- the block's root is
 *            its one member outside subb[b];
 *   blocks[31] bit b set when block b is a simple cycle (at least three
 *            members, each with exactly two neighbours inside the block).
 * While the DFS runs, subb[32], subb[31], ... hold the vertex stack: the
 * stack depth plus the blocks emitted never exceeds |S| - 1 <= 31. */
#define BSET_W 33
__device__ __forceinline__ int dfs_blocks(uint32_t S, const uint32_t *adj, uint32_t *dfs, int stride,
                                          uint32_t *blocks, uint32_t *own, uint32_t *subb) {
#define DFS(v) dfs[(v) * stride]
#define STK(i) subb[32 - (i)]
    int nblk = 0;
    const int root = __builtin_ctz(S);
    for (uint32_t x = S; x; x &= x - 1u) own[__builtin_ctz(x)] = 0u;
    uint32_t visited = 1u << root;
    uint32_t t = 1;
    int sp = 0;
    STK(sp++) = (uint32_t)root;
    /* du = DFS(u) and au = adj[u] & S stay in registers while at u */
    uint32_t du = 0u | (255u << 8) | (255u << 16);
    DFS(root) = du;
    uint32_t au = adj[root] & S;
    int u = root, cursor = -1;
    while (true) {
        uint32_t gt = cursor < 0 ? 0xffffffffu : (uint32_t)(0xffffffffull << (cursor + 1));
        uint32_t rem = au & gt;
        if (rem) {
            int v = __builtin_ctz(rem);
            if (!((visited >> v) & 1u)) {
                du = t | (255u << 8) | ((uint32_t)u << 16);
                DFS(v) = du;
                t++;
                visited |= 1u << v;
                STK(sp++) = (uint32_t)v;
                au = adj[v] & S;
                u = v;
                cursor = -1;
            } else {
                if ((int)(du >> 16) != v) {       /* not the tree edge to u's parent */
                    uint32_t dv = DFS(v);
                    if ((dv & 0xffu) < (du & 0xffu)) {
                        uint32_t low = (du >> 8) & 0xffu;
                        uint32_t nl = (dv & 0xffu) < low ? (dv & 0xffu) : low;
                        du = (du & 0xffff00ffu) | (nl << 8);
                        DFS(u) = du;
                    }
                }
                cursor = v;
            }
            continue;
        }
        if (u == root) break;
        int p = (int)(du >> 16);
        uint32_t dp = DFS(p);
        const uint32_t ap = adj[p] & S;
        uint32_t lowu = (du >> 8) & 0xffu;
        if (lowu >= (dp & 0xffu)) {
            uint32_t rest = 0, sub = 0, v;
            do {                                   /* pop down to u */
                v = STK(--sp);
                const uint32_t o = own[v] | (1u << v);
                own[v] = o;
                rest |= 1u << v;
                sub |= o;
            } while (v != (uint32_t)u);
            blocks[nblk] = rest | (1u << p);
            subb[nblk] = sub;
            nblk++;
            own[p] |= sub;
        }
        uint32_t lowp = (dp >> 8) & 0xffu;
        if (lowu < lowp) {
            dp = (dp & 0xffff00ffu) | (lowu << 8);
            DFS(p) = dp;
        }
        cursor = u;
        u = p;
        du = dp;
        au = ap;
    }
#undef STK
#undef DFS
    own[root] = S;
    uint32_t cyc = 0;
    for (int b = 0; b < nblk; b++) {
        const uint32_t blk = blocks[b];
        bool c = __builtin_popcount(blk) >= 3;
        for (uint32_t x = blk; c && x; x &= x - 1u) c = __builtin_popcount(adj[__builtin_ctz(x)] & blk) == 2;
        if (c) cyc |= 1u << b;
    }
    blocks[31] = cyc;
    return nblk;
}

/* The reference's grow(lowbit(x), S - (blk - x)) for a part x of block blk
 * (subb sb) that is connected inside the block: x's non-root members with
 * what hangs off them, plus, when x holds the block's root, everything
 * outside the block's subtree. */
__device__ __forceinline__ uint32_t block_side(uint32_t S, uint32_t x, uint32_t sb, const uint32_t *own) {
    uint32_t r = (x & ~sb) ? (S & ~sb) : 0u;
    for (uint32_t y = x & sb; y; y &= y - 1u) r |= own[__builtin_ctz(y)];
    return r;
}

__device__ __forceinline__ uint32_t pext32(uint32_t x, uint32_t mask) {
    uint32_t res = 0, bitpos = 1;
    while (mask) {
        uint32_t b = mask & (0u - mask);
        if (x & b) res |= bitpos;
        bitpos <<= 1;
        mask &= mask - 1;
    }
    return res;
}

/* Tree graphs (BlobHdr::is_tree): the split of connected S at the edge
 * (parent(c), c).  `sub` = the subtree of c inside S; o = 0 gives
 * (S - sub, sub), o = 1 gives (sub, S - sub).  In MPDP order that split has
 * rank 2 * post(c) + o, post(c) being c's position in the post-order of the
 * DFS from the lowest member of S with ascending neighbours (the block
 * emission order of find_blocks); in DPsub order its rank is pext(L, S). */
__device__ __forceinline__ uint32_t tree_post(uint32_t S, int c, const uint32_t *desc, const uint32_t *lglob) {
    return (uint32_t)(__builtin_popcount(S & desc[c]) - 1 + __builtin_popcount(S & lglob[c]));
}

/* Split `local` of a set whose blocks are blocks[0..nblk): block-major,
 * lb ascending over the block's non-empty proper subsets (lb = pdep of
 * local + 1 on the block); a valid split, so both sides are block_side. */
__device__ __forceinline__ void bicc_split(uint32_t S, const uint32_t *blocks, const uint32_t *own,
                                           const uint32_t *subb, uint32_t local, uint32_t *L, uint32_t *R) {
    uint32_t acc = 0, blk;
    int b = 0;
    while (true) {
        blk = blocks[b];
        uint32_t c = (1u << __builtin_popcount(blk)) - 2u;
        if (local < acc + c) break;
        acc += c;
        b++;
    }
    const uint32_t lb = pdep32(local - acc + 1u, blk), rb = blk ^ lb, sb = subb[b];
    *L = block_side(S, lb, sb, own);
    *R = block_side(S, rb, sb, own);
}

/* Enumerated splits of a block of m members: a simple cycle (bit b of
 * blocks[31]) has its m(m-1) arcs, any other block its 2^m - 2 non-empty
 * proper subsets. */
__device__ __forceinline__ uint32_t block_items(uint32_t blk, bool cyc) {
    const uint32_t m = (uint32_t)__builtin_popcount(blk);
    return cyc ? m * (m - 1u) : (1u << m) - 2u;
}

/* Enumerated splits of a set: *full = sum of 2^m - 2 over its blocks (the
 * reference's split count, which bounds the ranks), return = the splits
 * actually enumerated. */
__device__ __forceinline__ unsigned long long set_items(const uint32_t *blocks, int nblk,
                                                        unsigned long long *full) {
    unsigned long long n = 0, f = 0;
    for (int b = 0; b < nblk; b++) {
        const uint32_t m = (uint32_t)__builtin_popcount(blocks[b]);
        f += (1ull << m) - 2ull;
        n += ((blocks[31] >> b) & 1u) ? (unsigned long long)m * (m - 1u) : (1ull << m) - 2ull;
    }
    *full = f;
    return n;
}

/* Enumerated split `item` of a set with blocks[0..): (L, R) and its rank
 * `local` in bicc_split's order, which is the reference's.  A split lb of a
 * block is valid (the reference's L | R = S) exactly when lb and blk - lb
 * are both connected inside the block, a path between two members of a
 * block never leaving it; then L and R are block_side.  An invalid split
 * returns L = R = 0.  In a simple cycle only the arcs can be valid and they
 * are all that is enumerated: arc (start, len) runs len members along the
 * cycle from its start-th member, walking from the lowest member towards
 * its lower neighbour; its rank is pext(lb, blk) - 1 within the block. */
__device__ __forceinline__ uint32_t bicc_item(uint32_t S, const uint32_t *blocks, const uint32_t *own,
                                              const uint32_t *subb, uint32_t item, const uint32_t *adj,
                                              uint32_t *L, uint32_t *R) {
    const uint32_t cyc = blocks[31];
    uint32_t acc_i = 0, acc_r = 0, blk;
    bool is_cyc;
    int b = 0;
    while (true) {
        blk = blocks[b];
        is_cyc = (cyc >> b) & 1u;
        const uint32_t c = block_items(blk, is_cyc);
        if (item < acc_i + c) break;
        acc_i += c;
        acc_r += block_items(blk, false);
        b++;
    }
    const uint32_t local = item - acc_i;
    uint32_t lb, rank;
    if (is_cyc) {
        const uint32_t m1 = (uint32_t)__builtin_popcount(blk) - 1u;
        const uint32_t start = local / m1, len = local % m1 + 1u;
        uint32_t cur = blk & (0u - blk), prev = 0;
        lb = 0;
        for (uint32_t st = 0; st < start + len; st++) {
            if (st >= start) lb |= cur;
            uint32_t nb = adj[__builtin_ctz(cur)] & blk & ~prev;
            prev = cur;
            cur = nb & (0u - nb);
        }
        rank = acc_r + pext32(lb, blk) - 1u;
    } else {
        lb = pdep32(local + 1u, blk);
        rank = acc_r + local;
    }
    const uint32_t rb = blk ^ lb;
    if (!is_cyc && __builtin_popcount(blk) > 2 &&
        (grow(lb & (0u - lb), lb, adj) != lb || grow(rb & (0u - rb), rb, adj) != rb)) {
        *L = 0u;
        *R = 0u;
        return rank;
    }
    const uint32_t sb = subb[b];
    *L = block_side(S, lb, sb, own);
    *R = block_side(S, rb, sb, own);
    return rank;
}

/* One level, one kernel: a workgroup takes a batch of B sets.
 *   1. one lane per set prepares it (MPDP: block decomposition into LDS) and
 *      counts its splits; an exclusive scan lays the batch's splits out
 *      set-major, each set's splits in the reference's enumeration order;
 *   2. all lanes evaluate splits; per set, LDS atomics keep (a) the first
 *      valid split with its NaN flag and (b) the first minimum of the packed
 *      key (order-preserving cost bits << 32 | split rank) over non-NaN
 *      costs -- together the reference's "first found, replaced only if
 *      strictly cheaper" (gpuqo_cpu_common.cuh:189-211);
 *   3. the set's lane rebuilds the winning split and writes the memo.
 * DPsub splits are l = pdep(j, S), j = 1 .. 2^k-2 (l ascending,
 * gpuqo_cpu_dpsub.cu:32-79); only pairs whose sides are both connected are
 * costed, gathered first into an LDS queue so that every lane of a wave
 * costs a valid pair. */
/* sets of level k per wave of the tree evaluator: (k-1) * lpe lanes each */
__host__ __device__ constexpr uint32_t tree_spw(int k, int lpe = 1) { return WAVE / ((uint32_t)(k - 1) * lpe); }

/* LDS of the tree evaluator: the staged blob, then (one-lane form) a record
 * queue of WAVE entries per wave: slot, left, rows, startup, total, width */
#define TREE_Q_WORDS (6 * WAVE)
__host__ __device__ constexpr size_t tree_lds(int words, int lpe) {
    return (size_t)(((words + 3) & ~3) * 4) + (lpe == 1 ? (size_t)(BLOCK / WAVE) * TREE_Q_WORDS * 4 : 0);
}

/* Tree graphs, wave per set(s): a set of level k has k-1 edges and 2(k-1)
 * splits, (S - sub, sub) and (sub, S - sub) for the subtree `sub` below each
 * edge.  A segment of (k-1) * LPE consecutive lanes owns one set.  LPE 1:
 * lane e costs both splits of edge e, which share their two memo entries
 * (fewer lanes per split: the form for large levels); LPE 2: lanes 2e and
 * 2e+1 cost one split each (half the latency: the form for levels too
 * small to fill the GPU).  The segment reduces the first valid split (rank
 * and NaN flag) and the first minimum of the packed (cost, rank) key with
 * shuffles.  LPE 2: the winning lane writes its record itself, with the
 * plain width for a dense memo and with the hash table size packed for the
 * rank and open-addressing memos (whose big levels read these records).
 * LPE 1: the winners' records wait in the wave's LDS queue
 * (tree_lds) and are written a wave at a time, one record per lane, with
 * the hash table size packed (store_rel) -- one hash_pack pass per WAVE
 * sets instead of one per iteration.  Waves w0, w0 + wstride, ... cover the
 * level; every record is written before the call returns. */
/* LDS words of walk_tree_ws: set, left index, right index, pre-order
 * number and DFS stack (64 each) and a flag */
#define WALK_WS_WORDS (5 * 64 + 4)

/* Plan walk by one wave: breadth-first from root_set, one round per plan
 * level (the only dependent loads are left[S]), then lane 0 numbers the
 * nodes in pre-order and each lane writes one node.  ws: WALK_WS_WORDS of
 * LDS.  root_rec (optional): the root's left set and record {left, rows,
 * startup, total, width} when the calling wave has just computed them, so
 * that neither is read back from memory.  Returns the node count, or ~0u
 * for an inconsistent memo or too small a capacity. */
__device__ uint32_t walk_tree_ws(const MemoRef &mr, uint32_t root_set, DPNode *nodes, int cap, uint32_t *ws,
                                 const uint32_t *root_rec = nullptr) {
    uint32_t *s_set = ws;
    int *s_l = (int *)ws + 64, *s_r = (int *)ws + 128, *s_pre = (int *)ws + 192, *s_stack = (int *)ws + 256;
    volatile int *s_bad = (int *)ws + 320;
    const int lane = threadIdx.x & (WAVE - 1);
    if (lane == 0) {
        s_set[0] = root_set;
        *s_bad = cap < 1 || cap > 63;
    }
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    int f0 = 0, f1 = 1;
    while (f0 < f1 && !*s_bad) {
        const int i = f0 + lane;
        bool inner = false;
        uint32_t S = 0, l = 0;
        if (i < f1) {
            S = s_set[i];
            s_l[i] = -1;
            s_r[i] = -1;
            if (__builtin_popcount(S) > 1) {
                l = (i == 0 && root_rec) ? root_rec[0] : mr.left[memo_slot(mr, S)];
                if (l == 0 || (l & ~S) != 0 || l == S) *s_bad = 1;
                else inner = true;
            }
        }
        const unsigned long long m = __ballot(inner);
        const int newn = f1 + 2 * (int)__popcll(m);
        if (newn > cap) {
            if (lane == 0) *s_bad = 1;
        } else if (inner) {
            const int at = f1 + 2 * (int)__popcll(m & ((1ull << lane) - 1ull));
            s_set[at] = l;
            s_set[at + 1] = S ^ l;
            s_l[i] = at;
            s_r[i] = at + 1;
        }
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        f0 = f1;
        f1 = newn;
    }
    const int total = f0;
    if (*s_bad) return ~0u;
    if (lane == 0) {
        int sp = 0, cnt = 0;
        s_stack[sp++] = 0;
        while (sp > 0) {
            const int x = s_stack[--sp];
            s_pre[x] = cnt++;
            if (s_l[x] >= 0) {
                s_stack[sp++] = s_r[x];
                s_stack[sp++] = s_l[x];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane < total) {
        const uint32_t S = s_set[lane];
        DPNode nd;
        if (lane == 0 && root_rec) {
            nd.rows = __uint_as_float(root_rec[1]);
            nd.startup = __uint_as_float(root_rec[2]);
            nd.total = __uint_as_float(root_rec[3]);
            nd.width = (int)root_rec[4];
        } else {
            const float4 v = mr.vals[memo_slot(mr, S)];
            nd.rows = v.x; nd.startup = v.y; nd.total = v.z; nd.width = rec_width(__float_as_uint(v.w));
        }
        nd.set = S;
        nd.left = s_l[lane] >= 0 ? s_pre[s_l[lane]] : -1;
        nd.right = s_r[lane] >= 0 ? s_pre[s_r[lane]] : -1;
        nd.pad = 0;
        nodes[s_pre[lane]] = nd;
    }
    return (uint32_t)total;
}

/* walk_tree_ws with its scratch in static LDS (kernels of one wave) */
__device__ uint32_t walk_tree_wave(const MemoRef &mr, uint32_t root_set, DPNode *nodes, int cap) {
    __shared__ uint32_t ws[WALK_WS_WORDS];
    return walk_tree_ws(mr, root_set, nodes, cap, ws);
}

/* Level 2 of a tree DP needs no list and no memo loads: with every vertex
 * v > 0 having exactly one lower-numbered neighbour, its parent p(v), the
 * level-2 sets are the edges {p(v), v}, entry v - 1 of the implicit list,
 * and both sides of their one split are base relations, read from the
 * staged blob with the hash-table size packed as the memo holds it. */
__device__ __forceinline__ uint32_t edge_set(const uint32_t *adj, uint32_t i) {
    const uint32_t v = i + 1u;
    return (1u << __builtin_ctz(adj[v] & ((1u << v) - 1u))) | (1u << v);
}
__device__ __forceinline__ Rel base_rel(const BlobHdr &h, const uint32_t *blob, int i) {
    const float *b = (const float *)(blob + h.o_base) + 4 * i;
    Rel r;
    r.rows = b[0];
    r.startup = b[1];
    r.total = b[2];
    r.width = __float_as_int(b[3]);
    r.hs = hash_pack(r.rows, r.width, h.p.work_mem);
    return r;
}

template <int ECW, bool FAST, int LPE>
__device__ void eval_level_tree(const BlobHdr &h, uint32_t *smem, const View<uint32_t> &q, int k,
                                bool bicc, const uint32_t *lvl_sets, const uint32_t *lvl_slot, uint32_t count,
                                const MemoRef &mr, uint64_t w0, uint64_t wstride, uint32_t s0 = 0u,
                                bool implicit2 = false, uint32_t *root_rec = nullptr) {
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t per = (uint32_t)(k - 1) * LPE;
    const uint32_t spw = tree_spw(k, LPE);              /* sets per wave, per consecutive lanes each */
    const uint32_t seg = lane / per, t = lane % per, e = LPE == 1 ? t : t >> 1;
    const int o_lo = LPE == 1 ? 0 : (int)(t & 1u), o_hi = LPE == 1 ? 2 : o_lo + 1;
    const uint32_t *desc = smem + h.o_desc, *lglob = smem + h.o_lglob;
    /* rank memo: its tables in the staged blob */
    const uint32_t *rk = mr.rank ? smem + h.o_rank : nullptr;
    const int rn = h.n_rels;
    const uint64_t nwaves = (count + spw - 1) / spw;
    uint32_t *wq = smem + ((h.words + 3) & ~3) + (threadIdx.x / WAVE) * TREE_Q_WORDS;
    uint32_t qn = 0;                                    /* queued records (wave-uniform) */
    auto flush = [&]() {                                /* lane i writes queued record i */
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if ((uint32_t)lane < qn) {
            Rel r;
            r.rows = __uint_as_float(wq[2 * WAVE + lane]);
            r.startup = __uint_as_float(wq[3 * WAVE + lane]);
            r.total = __uint_as_float(wq[4 * WAVE + lane]);
            r.width = (int)wq[5 * WAVE + lane];
            const uint32_t slot = wq[lane];
            store_rel(mr.vals, slot, r, q.p.work_mem, true);
            mr.left[slot] = wq[WAVE + lane];
        }
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        qn = 0;
    };
    for (uint64_t w = w0; w < nwaves; w += wstride) {
        const uint64_t idx = w * spw + seg;
        const bool active = idx < count && seg < spw;
        /* per orientation o: first-valid key f_o, cost key k_o, result j_o
         * (scalars, not arrays: a loop-indexed local array goes to scratch) */
        uint32_t first = ~0u, f0 = ~0u, f1 = ~0u;
        unsigned long long key = ~0ull, k0 = ~0ull, k1 = ~0ull;
        Rel j0, j1;
        uint32_t S = 0, sub = 0, rS = 0;
        int c = 0;
        if (active) {
            S = implicit2 ? edge_set(smem + h.o_adj, (uint32_t)idx)
                          : ((w == w0 && s0) ? s0 : lvl_sets[idx]);   /* s0: this lane's first set, loaded by the caller */
            if (rk && lvl_slot) rS = lvl_slot[idx];     /* rank memo: the slot the expansion wrote */
            c = select_bit(S, e + 1u);                  /* member e+1: the root (lowest member) is skipped */
            sub = S & desc[c];
        }
        if (rk && !lvl_slot) {
            /* rank memo: every member but the lowest, t, is the lower end
             * of one edge, so the set's slot off[t] + sum Q[t][c] is a sum
             * over the segment (each edge counted once) */
            const int top = __builtin_ctz(S | 0x80000000u);
            uint32_t x = active && (LPE == 1 || (t & 1u) == 0u) ? rk[rn + top * rn + c] : 0u;
            for (uint32_t off = 1; off < per; off <<= 1) {
                const uint32_t y = __shfl_down(x, off);
                if (t + off < per) x += y;
            }
            const int hd = (int)(seg * per);
            x = __shfl(x, hd < WAVE ? hd : 0);
            rS = active ? rk[top] + x : 0u;
        }
        if (active) {
            const uint32_t rest = S ^ sub;
            uint32_t ss, sr;
            if (rk) {                                   /* slots of sub (lowest member c) and of S - sub */
                const uint32_t m1 = rank_digits(rk, rn, sub, c);   /* 1 + sig */
                ss = rk[c] + m1 - 1u;
                sr = rS - rk[rn + __builtin_ctz(S) * rn + c] * m1;
            } else {
                ss = memo_slot(mr, sub);
                sr = memo_slot(mr, rest);
            }
            const Rel Rs = implicit2 ? base_rel(h, smem, c) : load_rel(mr.vals, ss);
            const Rel Rr = implicit2 ? base_rel(h, smem, __builtin_ctz(S)) : load_rel(mr.vals, sr);
            uint32_t rk0, rk1;
            if (bicc) {
                rk0 = 2u * tree_post(S, c, desc, lglob);
                rk1 = rk0 + 1u;
            } else {                                    /* pext(S - sub, S) is the complement of pext(sub, S) */
                rk1 = pext32(sub, S);
                rk0 = rk1 ^ (0xffffffffu >> (32 - k));
            }
            uint64_t m[GQ_MAX_ECW];
            int nq = 0;
            const uint32_t *te = smem + h.o_tedge + 8 * c;
            if (!FAST) {
                const int nw = ECW ? ECW : h.ecw;
                if (h.tree_ecm) {                       /* the crossing classes of edge c, both orientations */
                    const uint64_t *em = (const uint64_t *)(smem + h.o_ecm) + c * nw;
                    for (int i = 0; i < nw; i++) {
                        m[i] = em[i];
                        nq += __popcll(m[i]);
                    }
                } else {
                    nq = match_mask<ECW>(q, rest, sub, m);
                }
            }
#pragma unroll 1
            for (int o = o_lo; o < o_hi; o++) {        /* o = 0: (S - sub, sub); o = 1: (sub, S - sub) */
                const uint32_t L = o ? sub : rest, R = o ? rest : sub;
                const Rel &RL = o ? Rs : Rr, &RR = o ? Rr : Rs;
                Rel r;
                if (FAST) {                             /* per-edge split table (BlobHdr::tree_fast) */
                    EdgeProv pv;
                    pv.sel = __uint_as_float(te[o]);
                    pv.nq = (int)(te[2] & 0xffu);
                    pv.n_mi = (int)((te[2] >> 8) & 0xffu);
                    pv.mi = (const uint16_t *)(te + 4 + 2 * o);
                    r = join_prov(q, L, RL, R, RR, pv);
                } else {
                    r = join_mask<ECW>(q, L, RL, R, RR, m, nq);
                }
                const bool nan = r.total != r.total;
                const uint32_t rk = o ? rk1 : rk0;
                const uint32_t fo = (rk << 1) | (nan ? 1u : 0u);
                const unsigned long long ko = nan ? ~0ull : (((unsigned long long)ord_key(r.total) << 32) | rk);
                if (o == 0) {
                    f0 = fo;
                    k0 = ko;
                    j0 = r;
                } else {
                    f1 = fo;
                    k1 = ko;
                    j1 = r;
                }
            }
            first = f0 < f1 ? f0 : f1;
            key = k0 < k1 ? k0 : k1;
        }
        /* segmented min over the set's lanes: after the step with offset
         * `off`, lane t holds the minimum over [t, min(t + 2 off, per)).
         * The first valid split's NaN flag decides only when some lane's
         * lower-ranked orientation costs NaN: otherwise no set's first split
         * has the flag, and that reduction is skipped (wave-uniform). */
        const int head = (int)(seg * per);               /* the set's first lane holds its minimum */
        uint32_t sf = 0u;
        if (__ballot(active && (first & 1u)) != 0ull) {
            sf = first;
            for (uint32_t off = 1; off < per; off <<= 1) {
                const uint32_t f2 = __shfl_down(sf, off);
                if (t + off < per) sf = f2 < sf ? f2 : sf;
            }
            sf = __shfl(sf, head < WAVE ? head : 0);
        }
        unsigned long long sk = key;
        for (uint32_t off = 1; off < per; off <<= 1) {
            const unsigned long long k2 = __shfl_down(sk, off);
            if (t + off < per) sk = k2 < sk ? k2 : sk;
        }
        sk = __shfl(sk, head < WAVE ? head : 0);
        bool win = false;
        uint32_t slot = 0, lft = 0;
        if (active) {
            const bool nanfirst = (sf & 1u) != 0;
            const bool w0n = nanfirst ? (f0 == sf) : (k0 == sk);
            const bool w1n = nanfirst ? (f1 == sf) : (k1 == sk);
            if (w0n || w1n) {
                win = true;
                slot = rk ? rS : (lvl_slot ? lvl_slot[idx] : S);
                lft = w1n ? sub : (S ^ sub);
                if (w1n) j0 = j1;
                if (LPE == 2) {
                    /* the plain width keeps hash_table_size off the
                     * latency-bound level's critical path; the large DPs
                     * of the rank and open-addressing memos pack it, since
                     * their big levels read these records (the served
                     * mix's 40 snowflakes of 25-30 relations: 39.3 against
                     * 34.7 ms packed; snowflake-20 dense: 0.141 against
                     * 0.146 ms unpacked) */
                    store_rel(mr.vals, slot, j0, q.p.work_mem, rk != nullptr || mr.keys != nullptr);
                    mr.left[slot] = lft;
                    if (root_rec) {                     /* one winning lane: the caller walks from this set */
                        root_rec[0] = lft;
                        root_rec[1] = __float_as_uint(j0.rows);
                        root_rec[2] = __float_as_uint(j0.startup);
                        root_rec[3] = __float_as_uint(j0.total);
                        root_rec[4] = (uint32_t)j0.width;
                    }
                }
            }
        }
        if (LPE == 1) {
            const unsigned long long bal = __ballot(win);
            if (win) {
                const uint32_t pos = qn + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                wq[pos] = slot;
                wq[WAVE + pos] = lft;
                wq[2 * WAVE + pos] = __float_as_uint(j0.rows);
                wq[3 * WAVE + pos] = __float_as_uint(j0.startup);
                wq[4 * WAVE + pos] = __float_as_uint(j0.total);
                wq[5 * WAVE + pos] = (uint32_t)j0.width;
            }
            qn += (uint32_t)__popcll(bal);
            if (qn + spw > WAVE) flush();
        }
    }
    if (LPE == 1 && qn) flush();
}

/* Tree graphs whose every vertex has its parent numbered below it
 * (BlobHdr::is_tree): the highest-numbered member m of a connected set T of
 * two or more relations is a leaf of T (its only possible neighbour in T is
 * its parent: its children are numbered above m), so it is T's
 * highest-numbered leaf, and removing it keeps T connected.  Every such T
 * is therefore generated exactly once, from S = T - max(T): the children of
 * connected S are S + v for the neighbours v of S numbered above max(S).
 * Returns that mask. */
__device__ __forceinline__ uint32_t top_leaf_children(const uint32_t *adj, uint32_t S) {
    uint32_t nb = 0;
    for (uint32_t r = S; r; r &= r - 1u) nb |= adj[__builtin_ctz(r)];
    const int m = 31 - __builtin_clz(S);
    return m == 31 ? 0u : nb & ~((2u << m) - 1u);
}

/* Dense- or rank-memo tree DP: the level list k+1 from set i of level k
 * (top_leaf_children), positions reserved with one atomic per wave on
 * *dst_count.  Called by whole waves; the order within a level list is
 * immaterial, each set's best split being computed on its own.  With
 * dst_slot (rank memo; rk = the rank tables), each new set's slot too. */
__device__ __forceinline__ void expand_level(const uint32_t *adj, const uint32_t *src, uint32_t count, uint32_t i,
                                             uint32_t *dst, uint32_t *dst_count, const uint32_t *src_slot = nullptr,
                                             uint32_t *dst_slot = nullptr, const uint32_t *rk = nullptr, int rn = 0) {
    const int lane = threadIdx.x & (WAVE - 1);
    uint32_t S = 0, emit = 0, sl = 0;
    if (i < count) {
        S = src ? src[i] : edge_set(adj, i);       /* src == nullptr: the implicit level-2 list */
        if (dst_slot) sl = src_slot[i];
        emit = top_leaf_children(adj, S);
    }
    const uint32_t cnt = (uint32_t)__builtin_popcount(emit);
    uint32_t incl = cnt;
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    const uint32_t total = __shfl(incl, WAVE - 1);
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(dst_count, total);
    uint32_t pos = __shfl(base, 0) + incl - cnt;
    /* rank memo: the new member v is above every member of S, so S + v has
     * S's top t and slot slot(S) + Q[t][v] */
    const uint32_t *q = rk ? rk + rn + __builtin_ctz(S | 0x80000000u) * rn : nullptr;
    for (; emit; emit &= emit - 1u, pos++) {
        const int v = __builtin_ctz(emit);
        dst[pos] = S | (1u << v);
        if (dst_slot) dst_slot[pos] = sl + q[v];
    }
}

/* k_eval_tree flags.  TREE_START: level 2 of a dense-memo DP whose lists
 * come by expansion -- its list is implicit (edge_set), its splits read the
 * base relations from the blob, and workgroup 0 stores the base relations
 * and counters[2] (the counters must be zero on entry).  TREE_WALK: the last
 * level (one set): workgroup 0 then walks the plan from it and writes the
 * 64 counters and the plan nodes straight to res_host (coherent pinned host
 * memory) -- words 56..57 and 58..59 being the device clock at the DP's
 * start and end -- then, after a system-scope release, word 48 (the walk's
 * status), and zeroes the device counters. */
#define TREE_START 1u
#define TREE_WALK 2u

/* Level k of a tree DP: workgroups [0, ev_blocks) evaluate it (the grid may
 * be sized by a bound on the level: idle ones leave after staging the
 * blob); with `next` set, workgroups from ev_blocks on build the level k+1
 * list there (dense or rank memo, counters[k+1] counts it; rank memo: the
 * slots go to next_slot), BLOCK level-k sets each.  The level's list has room for `cap` entries: each lane loads its
 * first set before the level size is known, so that this load, the level
 * size and the blob staging are in flight together. */
/* The one-lane per-edge-table form (the large levels) is bounded to five
 * waves per SIMD (at most 96 VGPRs, no spills): the compiler's own choice
 * can land just above 96 and drop it to four. */
template <int ECW, bool FAST, int LPE, uint32_t flags>
__global__ void __launch_bounds__(BLOCK, (FAST && LPE == 1) ? 5 : 1) k_eval_tree(BlobHdr h, const uint32_t *blob, int k, int bicc,
                                                      const uint32_t *lvl_sets, const uint32_t *lvl_slot,
                                                      MemoRef mr, uint32_t *counters, uint32_t ev_blocks,
                                                      uint32_t *next, uint32_t cap, uint32_t *next_slot,
                                                      uint32_t xs, uint32_t *res_host) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    /* xs > 1: the grid is xs times the level's workgroups, dispatched round
     * robin over the XCDs; only every xs-th works, so a small level runs on
     * one XCD and its writes stay in the L2 that the next level reads */
    uint32_t bid = blockIdx.x;
    if (xs > 1) {
        if (bid % xs) return;
        bid /= xs;
    }
    constexpr bool start = (flags & TREE_START) != 0;
    if (start && bid == 0 && threadIdx.x == 0) {     /* the DP's start on the device clock */
        const uint64_t t0 = wall_clock64();
        counters[56] = (uint32_t)t0;
        counters[57] = (uint32_t)(t0 >> 32);
    }
    if (bid >= ev_blocks) {                 /* adjacency (n words, inside the blob's LDS room) */
        const uint32_t count = start ? (uint32_t)(h.n_rels - 1) : counters[k];
        if ((int)threadIdx.x < h.n_rels) smem[threadIdx.x] = blob[h.o_adj + threadIdx.x];
        __syncthreads();
        expand_level(smem, start ? nullptr : lvl_sets, count, (bid - ev_blocks) * BLOCK + threadIdx.x, next,
                     counters + k + 1, lvl_slot, next_slot, mr.rank, mr.rn);
        return;
    }
    const uint32_t spw = tree_spw(k, LPE), per = (uint32_t)(k - 1) * LPE;
    const uint64_t w0 = (uint64_t)bid * (BLOCK / WAVE) + threadIdx.x / WAVE;
    const uint32_t seg = (threadIdx.x & (WAVE - 1)) / per;
    const uint64_t idx0 = w0 * spw + seg;
    const uint32_t count = start ? (uint32_t)(h.n_rels - 1) : counters[k];
    const uint32_t s_pre = (!start && seg < spw && idx0 < cap) ? lvl_sets[idx0] : 0u;
    for (int i = threadIdx.x; i < h.words; i += BLOCK) smem[i] = blob[i];
    __syncthreads();
    if constexpr (start) {
        if (bid == 0) {                     /* the base relations (read from the memo by later levels) */
        if ((int)threadIdx.x < h.n_rels) {
            store_base(mr.vals, 1u << threadIdx.x, h, smem, (int)threadIdx.x);
            mr.left[1u << threadIdx.x] = 0u;
        }
        if (threadIdx.x == 0) counters[2] = (uint32_t)(h.n_rels - 1);
        }
    }
    if ((uint64_t)bid * (BLOCK / WAVE) * spw >= count) return;
    View<uint32_t> q = make_view(h, smem);
    const uint32_t o_ws = (uint32_t)(tree_lds(h.words, LPE) / 4);   /* walk scratch after the queues */
    uint32_t *root_rec = (flags & TREE_WALK) ? smem + o_ws + WALK_WS_WORDS : nullptr;
    eval_level_tree<ECW, FAST, LPE>(h, smem, q, k, bicc != 0, lvl_sets, lvl_slot, count, mr, w0,
                                    (uint64_t)ev_blocks * (BLOCK / WAVE), idx0 < count ? s_pre : 0u, start,
                                    root_rec);
    if constexpr ((flags & TREE_WALK) != 0) {    /* one set, one workgroup: the plan walk */
        __syncthreads();
        if (threadIdx.x < WAVE) {
            const int n = h.n_rels, lane = (int)threadIdx.x;
            const uint32_t root = n == 32 ? 0xffffffffu : ((1u << n) - 1u);
            const uint32_t r = walk_tree_ws(mr, root, (DPNode *)(res_host + 64), 2 * n - 1, smem + o_ws, root_rec);
            const uint32_t cv = (start && lane == 2) ? (uint32_t)(n - 1) : counters[lane];
            const uint64_t t1 = wall_clock64();       /* the DP's end; its start is counters[56..57] */
            const uint32_t v = lane == 58 ? (uint32_t)t1 : lane == 59 ? (uint32_t)(t1 >> 32) : cv;
            if (lane != 48) res_host[lane] = v;
            counters[lane] = 0u;
            /* the status word last: a host polling it finds the nodes and
             * counters already in place */
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            __builtin_amdgcn_wave_barrier();
            if (lane == 48) __hip_atomic_store(res_host + 48, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    /* no pair counter: every split of a tree set is valid, so the host
     * counts sum_k |level k| * 2(k-1) -- one atomic per wave on a single
     * address would serialise the level */
}

/* Start of a dense- or rank-memo tree DP: zeroes the 64 result counters,
 * stores the base relations and builds level 2 (with slot2, the rank
 * memo's slots of its sets), the query's edges, each from its lower
 * end (top_leaf_children of a singleton).  One wave. */
__global__ void __launch_bounds__(WAVE) k_tree_start(BlobHdr h, const uint32_t *blob, float4 *vals, uint32_t *left,
                                                      uint32_t *lvl2, uint32_t *counters, uint32_t *slot2) {
    const int i = threadIdx.x;
    counters[i] = 0u;
    uint32_t emit = 0;
    if (i < h.n_rels) {
        const uint32_t slot = h.o_rank ? blob[h.o_rank + i] : (1u << i);   /* rank memo: off[i] */
        store_base(vals, slot, h, blob, i);
        left[slot] = 0u;
        emit = blob[h.o_adj + i] & ~((2u << i) - 1u);          /* neighbours above i */
    }
    const uint32_t cnt = (uint32_t)__builtin_popcount(emit);
    uint32_t incl = cnt;
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (i >= off) incl += y;
    }
    uint32_t pos = incl - cnt;
    for (; emit; emit &= emit - 1u, pos++) {
        const int j = __builtin_ctz(emit);
        lvl2[pos] = (1u << i) | (1u << j);
        if (slot2) slot2[pos] = blob[h.o_rank + i] + blob[h.o_rank + h.n_rels + i * h.n_rels + j];   /* off[i] + Q[i][j] */
    }
    __syncthreads();
    if (i == WAVE - 1) counters[2] = incl;
}

/* One DP's device state, as the batch kernels and the big-set kernels see it
 * (a single-query run is a batch of one).  No kernel writes a descriptor (they
 * write through its pointers only), so the evaluation kernels take the
 * array as __restrict__ and read its fields with scalar loads. */
struct BatchJob {
    BlobHdr h;
    const uint32_t *blob;
    float4 *vals;
    uint32_t *left, *conn, *lvl_sets, *counters;
    unsigned long long *lvl_best;
    uint32_t *lvl_first;
    DPNode *nodes;
    uint32_t lvl_off[33];
    uint32_t conn_words;
    int bicc;
    int cur_k;           /* step engine: the level this job evaluates next */
    /* open-addressing memo (tree graphs): keys, mask = capacity - 1, and the
     * slot of each level-list entry; keys == nullptr: dense memo */
    uint32_t *keys;
    uint32_t hmask;
    uint32_t *lvl_slot;
    uint32_t tree_iters;  /* wave-iterations of sets per tree task (Tuning::tree_iters when the job was made) */
    uint32_t expand;      /* dense-memo tree job whose level lists come from k_expand (no filter) */
};

__device__ __forceinline__ MemoRef job_memo(const BatchJob &J) {
    MemoRef m;
    m.vals = J.vals;
    m.left = J.left;
    m.keys = J.keys;
    m.mask = J.hmask;
    m.err = J.counters + 49;
    m.rank = J.h.o_rank ? J.blob + J.h.o_rank : nullptr;
    m.rn = J.h.n_rels;
    return m;
}

/* MPDP sets whose block decomposition yields more than BIG_SPLITS splits
 * (a biconnected block of 15 or more relations, e.g. a whole cycle) are not
 * evaluated by the workgroup that owns their batch: they are appended to a
 * list and their split range is spread over many workgroups, BIG_CHUNK
 * splits each (k_big_scan, k_big_eval). */
#define BIG_SPLITS 16384u
#define BIG_CHUNK 2048u

struct BigEnt {
    uint32_t job, gidx, cnt, done;   /* gidx: index into the job's level arrays */
};

struct BigList {
    BigEnt *ents;        /* nullptr: evaluate every set in place */
    uint32_t *count;
    uint32_t cap;
    uint32_t job;
    uint32_t gidx0;      /* level-array index of lvl_sets[0] of this call */
};

/* LDS layout of the level evaluation, after the staged blob (W4 words):
 *   s_best[BMAX] u64, s_off[BMAX+2] u64, s_first[BMAX], s_set[BMAX], s_misc[4]
 *   MPDP (general graphs), batches of at most BMAX_B sets: s_nblk[BMAX],
 *     s_blocks[BMAX_B][32], dfs[32][BMAX_B], s_own[BMAX_B][BSET_W],
 *     s_subb[BMAX_B][BSET_W] (dfs_blocks)
 *   DPsub: q_l, q_i, q_r [QCAP] each, then the staged connectivity bitmap. */
#define BMAX_B 64
/* byte offsets of the two 64-bit slot arrays (LDS atomics on them need
 * 8-byte alignment: the dynamic LDS base is 16-byte aligned, GQ_LDS_ALIGN) */
__host__ __device__ constexpr size_t eval_off_best(int words) { return (size_t)(((words + 3) & ~3) * 4); }
__host__ __device__ constexpr size_t eval_off_off(int words) { return eval_off_best(words) + BMAX * 8; }
__host__ __device__ constexpr size_t eval_lds_common(int words) {
    return eval_off_off(words) + (BMAX + 2) * 8 + BMAX * 4 * 2 + 16;
}
__host__ __device__ constexpr bool eval_slots_aligned(int words) {
    return eval_off_best(words) % 8 == 0 && eval_off_off(words) % 8 == 0 && eval_lds_common(words) % 4 == 0;
}
static_assert(eval_slots_aligned(1) && eval_slots_aligned(2) && eval_slots_aligned(3) && eval_slots_aligned(4) &&
                  eval_slots_aligned(5) && eval_slots_aligned(4093) && eval_slots_aligned(16383),
              "64-bit LDS slots of the level evaluator must be 8-byte aligned for every blob size");
static_assert(BMAX % 2 == 0, "s_off must start 8-byte aligned after s_best");
__host__ __device__ constexpr size_t eval_lds_bicc(int words) {
    return eval_lds_common(words) + BMAX * 4 + (size_t)BMAX_B * 32 * 4 * 2 + (size_t)BMAX_B * BSET_W * 4 * 2;
}
__host__ __device__ constexpr size_t eval_lds_dpsub(int words, int conn_words_lds) {
    return eval_lds_common(words) + 3 * QCAP * 4 + (size_t)conn_words_lds * 4;
}

/* Evaluate level k: batches bt0, bt0 + bstride, ... of B sets each (at most
 * BMAX_B for MPDP on general graphs).  The blob is staged at smem[0..words)
 * (and, for DPsub, `cb` points to the connectivity bitmap, staged or
 * global).  JRANK (with BICC, general graphs): the block enumeration gives
 * the valid splits and each is ranked by DPsub's j = pext(L, S), so the
 * result is DPsub's (dp_mode 2).  Returns the pairs costed; raises *err
 * when a set has more than 2^31 splits. */
template <bool BICC, bool JRANK = false>
__device__ uint32_t eval_level(const BlobHdr &h, uint32_t *smem, const View<uint32_t> &q, int k, uint32_t B,
                               const uint32_t *cb, const uint32_t *lvl_sets, uint32_t count, float4 *vals,
                               uint32_t *left, uint32_t bt0, uint32_t bstride, const BigList &bl) {
    if (BICC && h.is_tree == 0 && B > BMAX_B) B = BMAX_B;
    const bool tree = h.is_tree != 0;
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    unsigned long long *s_best = (unsigned long long *)((char *)smem + eval_off_best(h.words));
    unsigned long long *s_off = (unsigned long long *)((char *)smem + eval_off_off(h.words));   /* BMAX + 2 */
    uint32_t *s_first = (uint32_t *)(s_off + BMAX + 2);
    uint32_t *s_set = s_first + BMAX;
    uint32_t *s_misc = s_set + BMAX;                           /* [0] queue length, [1] error */
    uint32_t *s_nblk = s_misc + 4;
    uint32_t *s_blocks = s_nblk + BMAX;                        /* BICC: [BMAX_B][32]          */
    uint32_t *dfs = s_blocks + BMAX_B * 32;                    /* BICC: [32][BMAX_B]          */
    uint32_t *s_own = dfs + 32 * BMAX_B;                       /* BICC: [BMAX_B][BSET_W]      */
    uint32_t *s_subb = s_own + BMAX_B * BSET_W;                /* BICC: [BMAX_B][BSET_W]      */
    uint32_t *q_l = s_misc + 4;                                /* DPsub: [QCAP]               */
    uint32_t *q_i = q_l + QCAP;
    uint32_t *q_r = q_i + QCAP;
    const uint32_t *adj = smem + h.o_adj;
    const uint32_t T = (k >= 32) ? 0xfffffffeu : (uint32_t)((1ull << k) - 2ull);
    const uint32_t nbatches = (count + B - 1) / B;
    (void)lane; (void)s_nblk; (void)q_r;
    uint32_t npairs = 0;
    for (uint32_t bt = bt0; bt < nbatches; bt += bstride) {
        const uint32_t base = bt * B;
        const uint32_t nb = (count - base) < B ? (count - base) : B;
        if ((uint32_t)tid < nb) {
            uint32_t S = lvl_sets[base + tid];
            s_set[tid] = S;
            s_best[tid] = ~0ull;
            s_first[tid] = ~0u;
            unsigned long long cnt = tree ? 2ull * (k - 1) : T;
            if (BICC && !tree) {
                int nblk = dfs_blocks(S, adj, dfs + tid, BMAX_B, s_blocks + tid * 32, s_own + tid * BSET_W,
                                      s_subb + tid * BSET_W);
                s_nblk[tid] = nblk;
                unsigned long long full;
                cnt = set_items(s_blocks + tid * 32, nblk, &full);
                if (full > 0x7fffffffull || (JRANK && k > 31)) {   /* split rank would not fit 31 bits */
                    s_misc[1] = 1;
                    cnt = 0;
                } else if (bl.ents && cnt > BIG_SPLITS) {
                    const uint32_t slot = atomicAdd(bl.count, 1u);
                    if (slot < bl.cap) {
                        BigEnt e;
                        e.job = bl.job;
                        e.gidx = bl.gidx0 + base + (uint32_t)tid;
                        e.cnt = (uint32_t)cnt;
                        e.done = 0;
                        bl.ents[slot] = e;
                        cnt = 0;                    /* evaluated by k_big_eval */
                    }
                }
            }
            s_off[tid] = cnt;
        }
        __syncthreads();
        if (tid == 0) {
            unsigned long long acc = 0;
            for (uint32_t i = 0; i < nb; i++) {
                unsigned long long c = s_off[i];
                s_off[i] = acc;
                acc += c;
            }
            s_off[nb] = acc;
        }
        __syncthreads();
        const unsigned long long total = s_off[nb];
        if (tree) {
            const uint32_t per = 2u * (k - 1);
            const uint32_t *desc = smem + h.o_desc, *lglob = smem + h.o_lglob;
            for (unsigned long long item = tid; item < total; item += BLOCK) {
                const uint32_t i = (uint32_t)(item / per), local = (uint32_t)(item % per);
                const uint32_t S = s_set[i];
                const int c = select_bit(S, (local >> 1) + 1u);         /* skips the root (lowest) */
                const uint32_t sub = S & desc[c];
                const uint32_t o = local & 1u;
                const uint32_t L = o ? sub : (S ^ sub), R = S ^ L;
                const uint32_t rank = (BICC && !JRANK) ? (2u * tree_post(S, c, desc, lglob) + o) : pext32(L, S);
                Rel jr = join(q, L, load_rel(vals, L), R, load_rel(vals, R));
                npairs++;
                bool nan = jr.total != jr.total;
                atomicMin(&s_first[i], (rank << 1) | (nan ? 1u : 0u));
                if (!nan) atomicMin(&s_best[i], ((unsigned long long)ord_key(jr.total) << 32) | rank);
            }
        } else if (BICC) {
            for (unsigned long long item = tid; item < total; item += BLOCK) {
                uint32_t lo = 0, hi = nb;                      /* largest i with s_off[i] <= item */
                while (hi - lo > 1) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (s_off[mid] <= item) lo = mid;
                    else hi = mid;
                }
                const uint32_t i = lo, S = s_set[i];
                uint32_t L, R;
                const uint32_t brank = bicc_item(S, s_blocks + i * 32, s_own + i * BSET_W, s_subb + i * BSET_W,
                                                 (uint32_t)(item - s_off[i]), adj, &L, &R);
                if ((L | R) == S) {
                    const uint32_t rank = JRANK ? pext32(L, S) : brank;
                    Rel jr = join(q, L, load_rel(vals, L), R, load_rel(vals, R));
                    npairs++;
                    bool nan = jr.total != jr.total;
                    atomicMin(&s_first[i], (rank << 1) | (nan ? 1u : 0u));
                    if (!nan) atomicMin(&s_best[i], ((unsigned long long)ord_key(jr.total) << 32) | rank);
                }
            }
        } else {
            for (unsigned long long r0 = 0; r0 < total; r0 += QCAP) {
                if (tid == 0) s_misc[0] = 0;
                __syncthreads();
                unsigned long long rend = r0 + QCAP < total ? r0 + QCAP : total;
                for (unsigned long long item0 = r0; item0 < rend; item0 += BLOCK) {
                    unsigned long long item = item0 + tid;
                    bool ok = false;
                    uint32_t l = 0, i = 0, j = 0;
                    if (item < rend) {
                        i = (uint32_t)(item / T);
                        j = split_by_size((uint32_t)(item % T), k);
                        uint32_t S = s_set[i];
                        l = pdep32(j, S);
                        ok = is_conn(cb, l) && is_conn(cb, S ^ l);
                    }
                    unsigned long long m = __ballot(ok);
                    if (m) {
                        int leader = __builtin_ctzll(m);
                        uint32_t pos = 0;
                        if (lane == leader) pos = atomicAdd(&s_misc[0], (uint32_t)__popcll(m));
                        pos = __shfl(pos, leader);
                        if (ok) {
                            pos += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                            q_l[pos] = l;
                            q_i[pos] = i;
                            q_r[pos] = j;
                        }
                    }
                }
                __syncthreads();
                const uint32_t qn = s_misc[0];
                for (uint32_t e = tid; e < qn; e += BLOCK) {
                    uint32_t l = q_l[e], i = q_i[e], j = q_r[e];
                    uint32_t r = s_set[i] ^ l;
                    Rel jr = join(q, l, load_rel(vals, l), r, load_rel(vals, r));
                    npairs++;
                    bool nan = jr.total != jr.total;
                    atomicMin(&s_first[i], (j << 1) | (nan ? 1u : 0u));
                    if (!nan) atomicMin(&s_best[i], ((unsigned long long)ord_key(jr.total) << 32) | j);
                }
                __syncthreads();
            }
        }
        __syncthreads();
        if ((uint32_t)tid < nb) {
            uint32_t f = s_first[tid];
            if (f != ~0u) {
                uint32_t rank = (f & 1u) ? (f >> 1) : (uint32_t)(s_best[tid] & 0xffffffffull);
                uint32_t S = s_set[tid], L, R;
                if (tree && BICC && !JRANK) {
                    const uint32_t *desc = smem + h.o_desc, *lglob = smem + h.o_lglob;
                    uint32_t rest = S & (S - 1u), sub = 0;
                    while (rest) {
                        int c = __builtin_ctz(rest);
                        rest &= rest - 1u;
                        if (tree_post(S, c, desc, lglob) == (rank >> 1)) {
                            sub = S & desc[c];
                            break;
                        }
                    }
                    L = (rank & 1u) ? sub : (S ^ sub);
                    R = S ^ L;
                } else if (BICC && !JRANK) {
                    bicc_split(S, s_blocks + tid * 32, s_own + tid * BSET_W, s_subb + tid * BSET_W, rank, &L, &R);
                } else {
                    L = pdep32(rank, S);
                    R = S ^ L;
                }
                Rel jr = join(q, L, load_rel(vals, L), R, load_rel(vals, R));
                store_rel(vals, S, jr, q.p.work_mem);
                left[S] = L;
            }
        }
        __syncthreads();
    }
    return npairs;
}

/* Five waves per SIMD: at most 96 VGPRs, without spills (98-100 left alone,
 * four waves). */
template <bool BICC, bool JRANK = false>
__global__ void __launch_bounds__(BLOCK, 5) k_eval_wg(BlobHdr h, const uint32_t *blob, int k, uint32_t B,
                                                    int conn_words_lds, const uint32_t *lvl_sets, float4 *vals,
                                                    uint32_t *left, const uint32_t *conn, uint32_t *counters,
                                                    BigList bl) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    if ((uint64_t)blockIdx.x * B >= counters[k]) return;      /* no batch for this workgroup */
    const int tid = threadIdx.x;
    const int W4 = (h.words + 3) & ~3;
    uint32_t *s_misc = smem + W4 + BMAX * 2 + (BMAX + 2) * 2 + BMAX * 2;
    uint32_t *conn_l = s_misc + 4 + 3 * QCAP;
    for (int i = tid; i < h.words; i += BLOCK) smem[i] = blob[i];
    if (!BICC)
        for (int i = tid; i < conn_words_lds; i += BLOCK) conn_l[i] = conn[i];
    if (tid == 0) s_misc[1] = 0;
    __syncthreads();
    View<uint32_t> q = make_view(h, smem);
    const uint32_t *cb = conn_words_lds ? conn_l : conn;
    uint32_t npairs = eval_level<BICC, JRANK>(h, smem, q, k, B, cb, lvl_sets, counters[k], vals, left, blockIdx.x,
                                              gridDim.x, bl);
    if (tid == 0 && s_misc[1]) atomicOr(&counters[49], 1u);
    add_pairs(counters, npairs);
}

/* Exclusive scan of one value per thread over a workgroup of up to 1024
 * threads; `part` is 17 words of LDS.  Returns the thread's prefix and sets
 * *total. */
__device__ uint32_t block_scan_excl(uint32_t v, uint32_t *part, uint32_t *total) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    uint32_t x = v;
    for (int off = 1; off < WAVE; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == WAVE - 1) part[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int i = 0; i < nw; i++) {
            uint32_t t = part[i];
            part[i] = acc;
            acc += t;
        }
        part[16] = acc;
    }
    __syncthreads();
    const uint32_t r = x - v + part[w];
    *total = part[16];
    __syncthreads();
    return r;
}

/* Lays out the chunks of the deferred big sets: prefix[i] = first chunk of
 * entry i, prefix[nb] = total.  `hdr[0]` is the append counter (reset here
 * for the next level), `hdr[1]` receives the entry count. */
__global__ void __launch_bounds__(1024) k_big_scan(const BigEnt *ents, uint32_t *hdr, uint32_t cap,
                                                    uint32_t *prefix) {
    __shared__ uint32_t part[17];
    const uint32_t raw = hdr[0];
    const uint32_t nb = raw < cap ? raw : cap;
    const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = b0; i < b0 + per && i < nb; i++) sum += (ents[i].cnt + BIG_CHUNK - 1) / BIG_CHUNK;
    uint32_t total;
    uint32_t acc = block_scan_excl(sum, part, &total);
    for (uint32_t i = b0; i < b0 + per && i < nb; i++) {
        prefix[i] = acc;
        acc += (ents[i].cnt + BIG_CHUNK - 1) / BIG_CHUNK;
    }
    if (threadIdx.x == 0) {
        prefix[nb] = total;
        hdr[1] = nb;
        hdr[0] = 0;
    }
}

/* Big MPDP sets, chunk-parallel: workgroup-iteration c takes BIG_CHUNK
 * consecutive enumerated splits (bicc_item) of one deferred set (its block decomposition is
 * rebuilt in LDS when the set changes), keeps per lane the first valid
 * split (rank, NaN flag) and the first minimum of (cost, rank), reduces them
 * per wave and merges them into the set's lvl_first / lvl_best with global
 * atomics.  The workgroup finishing a set's last chunk (counted in
 * BigEnt::done) rebuilds the winning split and writes the memo. */
__global__ void __launch_bounds__(BLOCK) k_big_eval(const BatchJob *__restrict__ jobs, BigEnt *ents, const uint32_t *hdr,
                                                     const uint32_t *prefix) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    const uint32_t nb = hdr[1];
    if (nb == 0) return;
    const uint32_t total = prefix[nb];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    int have_job = -1;
    uint32_t have_b = ~0u;
    for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
        uint32_t lo = 0, hi = nb;                      /* largest b with prefix[b] <= c */
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (prefix[mid] <= c) lo = mid;
            else hi = mid;
        }
        const uint32_t b = lo;
        const uint32_t job = ents[b].job, gidx = ents[b].gidx, cnt = ents[b].cnt;
        const BatchJob &J = jobs[job];
        const BlobHdr &h = J.h;
        const int W4 = (h.words + 3) & ~3;
        uint32_t *s_blocks = smem + W4;                /* [32] */
        uint32_t *s_dfs = s_blocks + 32;               /* [32] */
        uint32_t *s_own = s_dfs + 32;                  /* [BSET_W] */
        uint32_t *s_subb = s_own + BSET_W;             /* [BSET_W] */
        uint32_t *s_misc = s_subb + BSET_W;            /* [0] nblk, [1] S, [2] last */
        if ((int)job != have_job) {
            __syncthreads();
            for (int i = tid; i < h.words; i += BLOCK) smem[i] = J.blob[i];
            __syncthreads();
            have_job = (int)job;
            have_b = ~0u;
        }
        const uint32_t *adj = smem + h.o_adj;
        if (b != have_b) {
            if (tid == 0) {
                const uint32_t S0 = J.lvl_sets[gidx];
                s_misc[1] = S0;
                s_misc[0] = (uint32_t)dfs_blocks(S0, adj, s_dfs, 1, s_blocks, s_own, s_subb);
            }
            __syncthreads();
            have_b = b;
        }
        const uint32_t S = s_misc[1];
        View<uint32_t> q = make_view(h, smem);
        const uint32_t c0 = (c - prefix[b]) * BIG_CHUNK;
        const uint32_t c1 = (cnt - c0) < BIG_CHUNK ? cnt : c0 + BIG_CHUNK;
        uint32_t first = ~0u, npairs = 0;
        unsigned long long best = ~0ull;
        for (uint32_t item = c0 + (uint32_t)tid; item < c1; item += BLOCK) {
            uint32_t L, R;
            const uint32_t rank = bicc_item(S, s_blocks, s_own, s_subb, item, adj, &L, &R);
            if ((L | R) == S) {
                Rel jr = join(q, L, load_rel(J.vals, L), R, load_rel(J.vals, R));
                npairs++;
                const bool nan = jr.total != jr.total;
                const uint32_t f = (rank << 1) | (nan ? 1u : 0u);
                if (f < first) first = f;
                if (!nan) {
                    const unsigned long long key = ((unsigned long long)ord_key(jr.total) << 32) | rank;
                    if (key < best) best = key;
                }
            }
        }
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            const uint32_t f2 = __shfl_xor(first, off);
            const unsigned long long b2 = __shfl_xor(best, off);
            const uint32_t p2 = __shfl_xor(npairs, off);
            first = f2 < first ? f2 : first;
            best = b2 < best ? b2 : best;
            npairs += p2;
        }
        if (lane == 0) {
            if (first != ~0u) atomicMin(&J.lvl_first[gidx], first);
            if (best != ~0ull) atomicMin(&J.lvl_best[gidx], best);
            if (npairs) atomicAdd((unsigned long long *)(J.counters + 40), (unsigned long long)npairs);
        }
        __threadfence();
        __syncthreads();
        if (tid == 0) {
            const uint32_t nch = (cnt + BIG_CHUNK - 1) / BIG_CHUNK;
            const uint32_t old = atomicAdd(&ents[b].done, 1u);
            s_misc[2] = (old == nch - 1) ? 1u : 0u;
        }
        __syncthreads();
        if (tid == 0 && s_misc[2]) {
            __threadfence();
            const uint32_t f = atomicOr(&J.lvl_first[gidx], 0u);
            if (f != ~0u) {
                const unsigned long long bst = atomicOr(&J.lvl_best[gidx], 0ull);
                const uint32_t rank = (f & 1u) ? (f >> 1) : (uint32_t)(bst & 0xffffffffull);
                uint32_t L, R;
                bicc_split(S, s_blocks, s_own, s_subb, rank, &L, &R);
                Rel jr = join(q, L, load_rel(J.vals, L), R, load_rel(J.vals, R));
                store_rel(J.vals, S, jr, q.p.work_mem);
                J.left[S] = L;
            }
        }
        __syncthreads();
    }
}

__host__ __device__ constexpr size_t big_lds(int words) {
    return (size_t)(((words + 3) & ~3) * 4) + (32 + 32 + 2 * BSET_W + 4) * 4;
}
/* ================================================================ batches
 *
 * Many DPs at once, level bucket by level bucket: one filter launch over the
 * candidates of every query, one planning launch that turns the level sizes
 * into per-level work tasks, then for each level k one launch evaluating the
 * level-k sets of all queries that have that level, and one plan-walk
 * launch.  A task is a slice of one query's level: its workgroup stages that
 * query's cost-model description in LDS and runs the same evaluators as the
 * single-query path. */



/* sets per task of level k: trees, else MPDP (bicc) or DPsub sets */
__host__ __device__ __forceinline__ uint32_t task_sets_of(bool is_tree, bool bicc, int k, uint32_t tree_iters) {
    if (is_tree) return (BLOCK / WAVE) * tree_spw(k) * tree_iters;
    uint64_t items = bicc ? 2ull * (k - 1) : ((1ull << k) - 2ull);
    /* MPDP: about one pass of the workgroup over the splits of its sets,
     * since each split is a long dependent chain; DPsub: three */
    uint64_t B = ((bicc ? BLOCK : 768) + items - 1) / items;
    const uint64_t bmax = bicc ? BMAX_B : BMAX;
    return (uint32_t)(B < 1 ? 1 : (B > bmax ? bmax : B));
}
#define RES_WORDS (64 + 64 * (sizeof(DPNode) / 4))   /* counters + up to 64 plan nodes per DP */

/* ---------------------------------------------------------------- allocation-free DPs
 * gq_plan_noalloc runs a dense-memo DPsub/MPDP DP of at most
 * GQ_NOALLOC_MAX_RELS relations entirely in `scratch`, a static device
 * array that every GPU the library runs on holds from the library's load
 * to the end of the process (NA_BYTES, about 40 MB).  Layout, 256-byte
 * aligned: vals 16 B x 2^20 | left 4 B x 2^20 | connectivity bitmap |
 * level lists 4 B x 2^20 | lvl_best 8 B x 2^20 | lvl_first 4 B x 2^20 |
 * counters + plan nodes (RES_WORDS) | staged blob and level offsets |
 * the big-set list (C(20,10) entries) and its prefix.  Host code never
 * uses the array's name as an address: ctx_init takes each device's
 * address once with hipGetSymbolAddress (g_scratch). */
namespace na {
constexpr size_t al(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NS = 1ull << GQ_NOALLOC_MAX_RELS;
constexpr size_t MAXC = 184756;                         /* C(20, 10): the largest level */
constexpr size_t STAGE_WORDS = 16384 + 36 + 32 + 128;   /* blob (64 KB LDS bound), offsets, level 1, job */
constexpr size_t O_VALS = 0;
constexpr size_t O_LEFT = O_VALS + al(16 * NS);
constexpr size_t O_CONN = O_LEFT + al(4 * NS);
constexpr size_t O_SETS = O_CONN + al(NS / 8);
constexpr size_t O_BEST = O_SETS + al(4 * NS);
constexpr size_t O_FIRST = O_BEST + al(8 * NS);
constexpr size_t O_RES = O_FIRST + al(4 * NS);
constexpr size_t O_BLOB = O_RES + al(RES_WORDS * 4);
constexpr size_t O_BIG = O_BLOB + al(STAGE_WORDS * 4);
constexpr size_t O_BIGP = O_BIG + al(MAXC * sizeof(BigEnt));
constexpr size_t BYTES = O_BIGP + al((MAXC + 1) * 4);
}  // namespace na
static_assert(na::BYTES < (64ull << 20), "the allocation-free scratch stays below 64 MB");
static __device__ __align__(256) unsigned char scratch[na::BYTES];

/* device address of `scratch` per device (0 until ctx_init runs there) */
#define GQ_MAX_DEVICES 64
static uintptr_t g_scratch[GQ_MAX_DEVICES];

__device__ __forceinline__ uint32_t task_sets(const BatchJob &J, int k) {
    return task_sets_of(J.h.is_tree != 0, J.bicc != 0, k, J.tree_iters);
}

/* Zeroes each job's 64 counters and stores its base relations (inserting
 * their keys first in an open-addressing memo, whose keys are cleared
 * beforehand).  With `blob_src`, job j's query blob is first copied from
 * blob_src + blob_at[j] (a staging area of this launch) into J.blob. */
__global__ void k_batch_init(const BatchJob *jobs, int nj, const uint32_t *blob_src, const uint32_t *blob_at) {
    for (int j = blockIdx.x; j < nj; j += gridDim.x) {
        const BatchJob &J = jobs[j];
        const int tid = threadIdx.x;
        const uint32_t *blob = J.blob;
        if (blob_src) {
            blob = blob_src + blob_at[j];
            for (uint32_t i = tid; i < (uint32_t)J.h.words; i += blockDim.x) const_cast<uint32_t *>(J.blob)[i] = blob[i];
        }
        if (tid < 64) J.counters[tid] = 0;
        __syncthreads();
        if (tid < J.h.n_rels) {
            const uint32_t s1 = 1u << tid;
            uint32_t slot = s1;
            if (J.keys) {                      /* keys cleared beforehand */
                bool fresh;
                slot = memo_insert(J.keys, J.hmask, s1, &fresh);
                if (slot == ~0u) atomicOr(&J.counters[49], 8u);
            } else if (J.h.o_rank) {           /* rank memo: off[tid] */
                slot = blob[J.h.o_rank + tid];
            }
            if (slot != ~0u) {
                store_base(J.vals, slot, J.h, blob, tid);
                J.left[slot] = 0;
            }
        }
    }
}

/* Copies n 16-byte words from host memory that the device can read (a
 * step's pinned packet) into device memory: the step stays on the compute
 * queue, with no copy-engine hand-off before its first kernel. */
__global__ void __launch_bounds__(BLOCK) k_stage(const uint4 *src, uint4 *dst, uint32_t n) {
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) dst[i] = src[i];
}

/* Clears the open-addressing keys of every job that has them. */
__global__ void __launch_bounds__(BLOCK) k_clear_keys(const BatchJob *jobs, int nj) {
    const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
    for (int j = 0; j < nj; j++) {
        uint32_t *keys = jobs[j].keys;
        if (!keys) continue;
        const uint64_t n = (uint64_t)jobs[j].hmask + 1u;
        for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) keys[i] = 0u;
    }
}

/* chunk_prefix[j] = first filter chunk of job j (host-computed) */
__global__ void __launch_bounds__(BLOCK) k_batch_filter(const BatchJob *jobs, int nj, const uint32_t *chunk_prefix) {
    __shared__ FilterSmem F;
    __shared__ int cur;
    const uint32_t total = chunk_prefix[nj];
    int have = -1;
    for (uint32_t c = blockIdx.x; c < total; c += gridDim.x) {
        if (threadIdx.x == 0) {
            int lo = 0, hi = nj;
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (chunk_prefix[mid] <= c) lo = mid;
                else hi = mid;
            }
            cur = lo;
        }
        __syncthreads();
        const int j = cur;
        const BatchJob &J = jobs[j];
        if (j != have) {
            filter_setup(F, J.h, J.blob);
            have = j;
        }
        filter_chunk(F, J.h, c - chunk_prefix[j], J.lvl_off, J.lvl_sets, J.lvl_best, J.lvl_first, J.conn,
                     J.conn_words, J.counters);
    }
}

/* Job classes of a level launch: tree jobs with the per-edge split table
 * run in k_batch_eval_tree (few registers, high occupancy), every other job
 * in k_batch_eval. */
#define JOBS_TREE 1
#define JOBS_GENERAL 2

__host__ __device__ __forceinline__ bool fast_tree_job(const BlobHdr &h) { return h.is_tree && h.tree_fast; }

__device__ __forceinline__ uint32_t job_tasks(const BatchJob &J, int k, int cls) {
    if (k < 2 || k > J.h.n_rels) return 0;
    if ((cls == JOBS_TREE) != fast_tree_job(J.h)) return 0;
    const uint32_t spt = task_sets(J, k);
    return (J.counters[k] + spt - 1) / spt;
}

/* task_prefix[k * (nj + 1) + j] = first task of job j at level k; one
 * workgroup per level k = 2 + blockIdx.x. */
__global__ void __launch_bounds__(1024) k_batch_plan(const BatchJob *jobs, int nj, uint32_t *task_prefix, int cls) {
    __shared__ uint32_t part[17];
    const int k = 2 + (int)blockIdx.x;
    uint32_t *tp = task_prefix + (size_t)k * (nj + 1);
    const uint32_t per = ((uint32_t)nj + blockDim.x - 1) / blockDim.x;
    const uint32_t j0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t j = j0; j < j0 + per && j < (uint32_t)nj; j++) sum += job_tasks(jobs[j], k, cls);
    uint32_t total;
    uint32_t acc = block_scan_excl(sum, part, &total);
    for (uint32_t j = j0; j < j0 + per && j < (uint32_t)nj; j++) {
        tp[j] = acc;
        acc += job_tasks(jobs[j], k, cls);
    }
    if (threadIdx.x == 0) tp[nj] = total;
}

/* Step engine: task_prefix[j] = first task of job j at its own level
 * J.cur_k (jobs past their last level have none); one workgroup. */
__global__ void __launch_bounds__(1024) k_step_plan(const BatchJob *jobs, int nj, uint32_t *task_prefix, int cls) {
    __shared__ uint32_t part[17];
    const uint32_t per = ((uint32_t)nj + blockDim.x - 1) / blockDim.x;
    const uint32_t j0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t j = j0; j < j0 + per && j < (uint32_t)nj; j++) sum += job_tasks(jobs[j], jobs[j].cur_k, cls);
    uint32_t total;
    uint32_t acc = block_scan_excl(sum, part, &total);
    for (uint32_t j = j0; j < j0 + per && j < (uint32_t)nj; j++) {
        task_prefix[j] = acc;
        acc += job_tasks(jobs[j], jobs[j].cur_k, cls);
    }
    if (threadIdx.x == 0) task_prefix[nj] = total;
}

/* One DP level of every job: level k_fixed of all of them (a batch run in
 * lockstep), or, with k_fixed = 0, each job's own level J.cur_k (the step
 * engine).  `tpg` is the task prefix of this launch; it is staged in LDS at
 * word `tp_lds` (0: read from global memory).  Non-tree MPDP sets with more
 * than BIG_SPLITS splits go to the big-set list (`big`, `big_hdr`). */
__global__ void __launch_bounds__(BLOCK) k_batch_eval(const BatchJob *__restrict__ jobs, int nj, int k_fixed,
                                                       const uint32_t *__restrict__ tpg, uint32_t tp_lds, BigEnt *big,
                                                       uint32_t big_cap, uint32_t *big_hdr) {
    /* The dynamic LDS must be the kernel's only LDS: a static __shared__
     * variable would move its base off the 8-byte alignment that the 64-bit
     * LDS atomics of eval_level need.  Each thread therefore finds its job
     * itself (the search is uniform across the workgroup). */
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    const uint32_t total = tpg[nj];
    if (blockIdx.x >= total) return;
    const uint32_t *tp = tpg;
    if (tp_lds) {
        uint32_t *tl = smem + tp_lds;
        for (int i = threadIdx.x; i <= nj; i += BLOCK) tl[i] = tpg[i];
        __syncthreads();
        tp = tl;
    }
    int have = -1;
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = nj;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (tp[mid] <= t) lo = mid;
            else hi = mid;
        }
        const int j = lo;
        const BatchJob &J = jobs[j];
        const BlobHdr &h = J.h;
        const int k = k_fixed > 0 ? k_fixed : J.cur_k;
        if (j != have) {
            __syncthreads();
            for (int i = threadIdx.x; i < h.words; i += BLOCK) smem[i] = J.blob[i];
            if (threadIdx.x == 0) smem[((h.words + 3) & ~3) + BMAX * 2 + (BMAX + 2) * 2 + BMAX * 2 + 1] = 0;
            __syncthreads();
            have = j;
        }
        View<uint32_t> q = make_view(h, smem);
        const uint32_t spt = task_sets(J, k);
        const uint32_t first = (t - tp[j]) * spt, cnt = J.counters[k];
        const uint32_t m = cnt - first < spt ? cnt - first : spt;
        const uint32_t *sets = J.lvl_sets + J.lvl_off[k] + first;
        uint32_t np;
        if (h.is_tree) {                       /* trees without the per-edge table */
            const MemoRef mr = job_memo(J);
            const uint32_t *slots = J.lvl_slot ? J.lvl_slot + J.lvl_off[k] + first : nullptr;
            if (h.ecw == 1)
                eval_level_tree<1, false, 1>(h, smem, q, k, J.bicc == 1, sets, slots, m, mr, threadIdx.x / WAVE,
                                                BLOCK / WAVE);
            else
                eval_level_tree<0, false, 1>(h, smem, q, k, J.bicc == 1, sets, slots, m, mr, threadIdx.x / WAVE,
                                                BLOCK / WAVE);
            np = 0;                            /* counted on the host */
        } else if (J.bicc == 2) {                /* DPsub splits through the blocks, evaluated in place */
            BigList bl;
            bl.ents = nullptr;
            bl.count = nullptr;
            bl.cap = 0;
            bl.job = 0;
            bl.gidx0 = 0;
            np = eval_level<true, true>(h, smem, q, k, spt, J.conn, sets, m, J.vals, J.left, 0, 1, bl);
        } else if (J.bicc) {
            BigList bl;
            bl.ents = big;
            bl.count = big_hdr;
            bl.cap = big_cap;
            bl.job = (uint32_t)j;
            bl.gidx0 = J.lvl_off[k] + first;
            np = eval_level<true>(h, smem, q, k, spt, J.conn, sets, m, J.vals, J.left, 0, 1, bl);
        } else {
            BigList bl;
            bl.ents = nullptr;
            bl.count = nullptr;
            bl.cap = 0;
            bl.job = 0;
            bl.gidx0 = 0;
            np = eval_level<false>(h, smem, q, k, spt, J.conn, sets, m, J.vals, J.left, 0, 1, bl);
        }
        __syncthreads();
        if (threadIdx.x == 0 && !h.is_tree) {
            /* the split-count flag of eval_level; the tree evaluator sets no
             * flag and its per-wave record queues reuse this LDS word */
            uint32_t *misc = smem + ((h.words + 3) & ~3) + BMAX * 2 + (BMAX + 2) * 2 + BMAX * 2;
            if (misc[1]) atomicOr(&J.counters[49], 1u);
        }
        add_pairs(J.counters, np);
    }
}

/* The tree jobs with a per-edge split table (k_batch_eval's task scheme;
 * k_batch_plan / k_step_plan with JOBS_TREE), one task per workgroup: the
 * grid is the exact task count, which the host derives from the exact level
 * sizes of tree queries (a device total above the grid would leave tasks
 * undone and raises error bit 32 of job 0).  A kernel of its own with only
 * this evaluator and no task loop keeps few registers and many waves per
 * SIMD: five (at most 96 VGPRs, no spills).  The job descriptors are
 * read-only here (__restrict__), so the job's header, memo and list
 * pointers are scalar loads kept in SGPRs; without that the compiler keeps
 * them in VGPRs (135 of them, three waves per SIMD, and the served mix
 * plans 20% slower).  Pairs are not counted: every split of a tree set is
 * valid. */
__global__ void __launch_bounds__(BLOCK, 5) k_batch_eval_tree(const BatchJob *__restrict__ jobs, int nj,
                                                               int k_fixed, const uint32_t *__restrict__ tpg) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    const uint32_t t = blockIdx.x;
    const uint32_t total = tpg[nj];
    if (t == 0 && total > gridDim.x) atomicOr(&jobs[0].counters[49], 32u);
    if (t >= total) return;
    int lo = 0, hi = nj;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (tpg[mid] <= t) lo = mid;
        else hi = mid;
    }
    const BatchJob &J = jobs[lo];
    const BlobHdr &h = J.h;
    const int k = k_fixed > 0 ? k_fixed : J.cur_k;
    for (int i = threadIdx.x; i < h.words; i += BLOCK) smem[i] = J.blob[i];
    __syncthreads();
    View<uint32_t> q = make_view(h, smem);
    const uint32_t spt = task_sets(J, k);
    const uint32_t first = (t - tpg[lo]) * spt, cnt = J.counters[k];
    /* the task prefix may come from the host's level sizes: a level that
     * came out shorter (an expansion error, flagged in counters[49]) leaves
     * tasks past its end */
    if (cnt <= first) return;
    const uint32_t m = cnt - first < spt ? cnt - first : spt;
    const MemoRef mr = job_memo(J);
    eval_level_tree<1, true, 1>(h, smem, q, k, J.bicc == 1, J.lvl_sets + J.lvl_off[k] + first,
                                   J.lvl_slot ? J.lvl_slot + J.lvl_off[k] + first : nullptr, m, mr,
                                   threadIdx.x / WAVE, BLOCK / WAVE);
}

/* Plan walk from `root_set` (dpsub_buildQueryTree, gpuqo_query_tree.cuh:46-72)
 * by one wave: the tree is expanded breadth-first, one round per tree level
 * (the only dependent loads are left[S] of the round's inner nodes), then
 * lane 0 numbers the nodes in pre-order -- node, left subtree, right
 * subtree, the order of the reference's tree -- and the lanes write the
 * nodes with their values.  Returns the node count, or ~0 when the memo is
 * inconsistent (a left side that is empty, not a proper subset, or a tree of
 * more than `cap` nodes).  cap <= 63.  Called by all 64 lanes of a
 * workgroup of one wave. */
/* Plan walk of finished jobs, one wave each: job jobs[idx ? idx[b] : b] for
 * workgroup b < nj.  res == nullptr: the nodes go to J.nodes and the status
 * to J.counters[48].  Otherwise workgroup b writes RES_WORDS words at
 * res + b * RES_WORDS: the job's 64 counters (status in word 48), then the
 * nodes, so one copy returns every result of the launch. */
__global__ void __launch_bounds__(WAVE) k_batch_tree(const BatchJob *jobs, int nj, const uint32_t *idx,
                                                      uint32_t *res) {
    const int b = blockIdx.x;
    if (b >= nj) return;
    const BatchJob &J = jobs[idx ? idx[b] : (uint32_t)b];
    const int n = J.h.n_rels;
    const uint32_t full = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
    uint32_t *out = res ? res + (size_t)b * RES_WORDS : nullptr;
    const uint32_t r = walk_tree_wave(job_memo(J), full, out ? (DPNode *)(out + 64) : J.nodes, 2 * n - 1);
    if (out) {
        if (threadIdx.x < 64) out[threadIdx.x] = threadIdx.x == 48 ? r : J.counters[threadIdx.x];
    } else if (threadIdx.x == 0) {
        J.counters[48] = r;
    }
}

/* Fills n words with v (clears an open-addressing key array). */
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

/* Open-addressing and rank memos, level list k from level k-1 (tree
 * graphs).  Every connected set T of k >= 2 relations is generated exactly
 * once, from S = T - (T's highest-numbered leaf) (top_leaf_children).  Each
 * thread takes one set S of level k-1 (the singletons for k = 2); the
 * workgroup reserves its list positions with one atomic, then, with the
 * open-addressing memo, each emitted set claims its key (a key already
 * present would be a duplicate: error bit 16).  Level k is level k_fixed of
 * every job, or each job's J.cur_k when k_fixed = 0; `prefix[j]` = first
 * task (BLOCK source sets) of job j, host-computed from the exact level
 * sizes; jobs with neither keys nor rank tables have no tasks, except
 * dense-memo tree jobs marked `expand` (their slot is the set itself, so
 * only the list is written).  A task belongs to one job, so the whole
 * workgroup works for it.
 * Order within a level list is immaterial: each set's best split is
 * computed independently. */
__global__ void __launch_bounds__(BLOCK) k_expand(const BatchJob *__restrict__ jobs, int nj, int k_fixed,
                                                   const uint32_t *__restrict__ prefix) {
    __shared__ uint32_t s_adj[32];
    __shared__ uint32_t part[17];
    const uint32_t total = prefix[nj];
    for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = nj;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (prefix[mid] <= t) lo = mid;
            else hi = mid;
        }
        const BatchJob &J = jobs[lo];
        const int k = k_fixed > 0 ? k_fixed : J.cur_k;
        const int n = J.h.n_rels;
        const uint32_t src_count = (k - 1 == 1) ? (uint32_t)n : J.counters[k - 1];
        const uint32_t i = (t - prefix[lo]) * BLOCK + threadIdx.x;
        const bool have = i < src_count && (J.keys != nullptr || J.h.o_rank != 0 || J.expand != 0);
        const uint32_t S = have ? ((k - 1 == 1) ? (1u << i) : J.lvl_sets[J.lvl_off[k - 1] + i]) : 0u;
        /* rank memo: the source's slot (off[i] for a singleton); each child
         * S + v, v above every member, has slot slot(S) + Q[top(S)][v] */
        const uint32_t *rq = J.h.o_rank ? J.blob + J.h.o_rank : nullptr;
        const uint32_t sl = (have && rq) ? ((k - 1 == 1) ? rq[i] : J.lvl_slot[J.lvl_off[k - 1] + i]) : 0u;
        if (threadIdx.x < 32) s_adj[threadIdx.x] = (int)threadIdx.x < n ? J.blob[J.h.o_adj + threadIdx.x] : 0u;
        __syncthreads();
        uint32_t emit = have ? top_leaf_children(s_adj, S) : 0u;
        const uint32_t cnt = (uint32_t)__builtin_popcount(emit);
        uint32_t all;
        const uint32_t off = block_scan_excl(cnt, part, &all);
        if (threadIdx.x == 0) part[0] = all ? atomicAdd(&J.counters[k], all) : 0u;
        __syncthreads();
        uint32_t pos = part[0] + off;
        __syncthreads();                           /* part and s_adj are rewritten by the next task */
        while (emit) {
            const int v = __builtin_ctz(emit);
            emit &= emit - 1u;
            const uint32_t T = S | (1u << v);
            if (J.keys) {
                bool fresh;
                const uint32_t slot = memo_insert(J.keys, J.hmask, T, &fresh);
                if (slot == ~0u || !fresh) {
                    atomicOr(&J.counters[49], slot == ~0u ? 8u : 16u);
                    continue;
                }
                J.lvl_slot[J.lvl_off[k] + pos] = slot;
            } else if (rq) {
                J.lvl_slot[J.lvl_off[k] + pos] = sl + rq[J.h.n_rels + __builtin_ctz(S) * J.h.n_rels + v];
            }                                      /* dense memo: the slot is the set */
            J.lvl_sets[J.lvl_off[k] + pos] = T;
            pos++;
        }
    }
}

/* ================================================================ DPsize
 *
 * gpuqo_algorithm = dpsize against the oracle cpu_dpsize
 * (gpuqo_cpu_dpsize.cu:27-93).  Level lists keep the CPU's creation order:
 * level 1 in the order the host supplies (libstdc++ unordered_map iteration
 * over the base relations, :43-47), level k in the order the sets are first
 * created by the pair scan
 *     for t = 0..k-2 (left size k-1-t):  for a in list[k-1-t]:  for b in list[t+1]
 * (:50-73).  The scan runs tile-parallel; each valid pair carries its global
 * rank (t, a, b).  Per set, atomics keep the first valid rank (with a NaN
 * flag) and the first cheapest pair; the first valid ranks then order the
 * new list: sets are bucketed by the row (t, a) of their first pair, rows
 * are laid out by an exclusive scan, and each row is sorted by b. */

#define DS_TA 64
#define DS_TB 64
#define DS_RANK_BITS 27

struct DsBufs {
    unsigned long long *best, *first;   /* [2^n]                                   */
    uint32_t *lset, *lnbr;              /* level lists, all levels concatenated    */
    uint32_t *raw, *tmp;                /* [maxC] new sets of the level            */
    uint32_t *rowcnt, *rowfill, *rowoff;/* [2^n (+1)]                              */
    uint32_t *sz;                       /* [0..32] size, [40..72] offset, [80..112] rowbase,
                                           [120..152] tile prefix, [160] rows, [161] tiles */
};

__global__ void k_ds_init(BlobHdr h, const uint32_t *blob, const int *level1, DsBufs d, uint64_t nsets,
                          float4 *vals, uint32_t *left, uint32_t *counters) {
    uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = tid; i < nsets; i += (uint64_t)gridDim.x * blockDim.x) {
        d.best[i] = ~0ull;
        d.first[i] = ~0ull;
    }
    if (tid < 64) counters[tid] = 0;
    if (tid < 256) d.sz[tid] = 0;
    if (tid < (uint64_t)h.n_rels) {
        uint32_t s1 = 1u << tid;
        store_base(vals, s1, h, blob, (int)tid);
        left[s1] = 0;
        int r = level1[tid];
        d.lset[tid] = 1u << r;
        d.lnbr[tid] = blob[h.o_adj + r];
    }
}

/* Sizes, offsets, row bases and tile counts of level k (one thread). */
__global__ void k_ds_prep(int k, int n, DsBufs d) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t *sz = d.sz;
    if (k == 2) { sz[1] = (uint32_t)n; sz[40 + 1] = 0; }
    sz[40 + k] = sz[40 + k - 1] + sz[k - 1];
    uint32_t rows = 0, tiles = 0;
    for (int t = 0; t <= k - 2; t++) {
        int ls = k - 1 - t, rs = t + 1;
        sz[80 + t] = rows;
        rows += sz[ls];
        sz[120 + t] = tiles;
        tiles += ((sz[ls] + DS_TA - 1) / DS_TA) * ((sz[rs] + DS_TB - 1) / DS_TB);
    }
    sz[120 + k - 1] = tiles;
    sz[160] = rows;
    sz[161] = tiles;
}

__global__ void k_ds_zero_rows(DsBufs d) {
    uint32_t rows = d.sz[160];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= rows; i += gridDim.x * blockDim.x) {
        d.rowcnt[i] = 0;
        d.rowfill[i] = 0;
    }
}

/* The pair scan of one DPsize level over (left list x right list) tiles.
 * Four waves per SIMD (at most 128 VGPRs, 8 bytes spilled; left alone the
 * compiler takes 135 and three waves: clique-18 DPsize 71.3 against
 * 88.7 ms; five waves spill 136 bytes). */
__global__ void __launch_bounds__(BLOCK, 4) k_ds_pairs(BlobHdr h, const uint32_t *blob, int k, DsBufs d,
                                                     const float4 *vals, uint32_t *counters) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const int W4 = (h.words + 3) & ~3;
    uint32_t *tl = smem + W4;                 /* left sets  [DS_TA] */
    uint32_t *tn = tl + DS_TA;                /* left nbrs  [DS_TA] */
    uint32_t *tr = tn + DS_TA;                /* right sets [DS_TB] */
    uint32_t *qq = tr + DS_TB;                /* queue [DS_TA*DS_TB]: a | b << 16 */
    uint32_t *qn = qq + DS_TA * DS_TB;
    for (int i = tid; i < h.words; i += BLOCK) smem[i] = blob[i];
    __syncthreads();
    View<uint32_t> q = make_view(h, smem);
    const uint32_t *sz = d.sz;
    const uint32_t tiles = sz[161];
    uint32_t checked = 0, npairs = 0;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        int t = 0;
        while ((uint32_t)sz[120 + t + 1] <= tile) t++;
        const int ls = k - 1 - t, rs = t + 1;
        const uint32_t nl = sz[ls], nr = sz[rs], ol = sz[40 + ls], orr = sz[40 + rs];
        const uint32_t ntb = (nr + DS_TB - 1) / DS_TB;
        const uint32_t local = tile - sz[120 + t];
        const uint32_t a0 = (local / ntb) * DS_TA, b0 = (local % ntb) * DS_TB;
        const uint32_t na = nl - a0 < DS_TA ? nl - a0 : DS_TA, nb = nr - b0 < DS_TB ? nr - b0 : DS_TB;
        if (tid < DS_TA) {
            tl[tid] = tid < (int)na ? d.lset[ol + a0 + tid] : 0u;
            tn[tid] = tid < (int)na ? d.lnbr[ol + a0 + tid] : 0u;
        } else if (tid < DS_TA + DS_TB) {
            int j = tid - DS_TA;
            tr[j] = j < (int)nb ? d.lset[orr + b0 + j] : 0u;
        }
        if (tid == 0) *qn = 0;
        __syncthreads();
        for (int p0 = 0; p0 < DS_TA * DS_TB; p0 += BLOCK) {
            int p = p0 + tid, a = p / DS_TB, b = p % DS_TB;
            bool ok = false;
            if (a < (int)na && b < (int)nb) {
                checked++;
                uint32_t l = tl[a], r = tr[b];
                ok = (l & r) == 0u && (tn[a] & r) != 0u;
            }
            unsigned long long m = __ballot(ok);
            if (m) {
                int leader = __builtin_ctzll(m);
                uint32_t pos = 0;
                if (lane == leader) pos = atomicAdd(qn, (uint32_t)__popcll(m));
                pos = __shfl(pos, leader);
                if (ok) qq[pos + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)a | ((uint32_t)b << 16);
            }
        }
        __syncthreads();
        const uint32_t nq = *qn;
        for (uint32_t e = tid; e < nq; e += BLOCK) {
            uint32_t a = qq[e] & 0xffffu, b = qq[e] >> 16;
            uint32_t l = tl[a], r = tr[b], S = l | r;
            uint32_t ag = a0 + a, bg = b0 + b;
            Rel jr = join(q, l, load_rel(vals, l), r, load_rel(vals, r));
            npairs++;
            bool nan = jr.total != jr.total;
            unsigned long long g = ((unsigned long long)t << (2 * DS_RANK_BITS)) |
                                   ((unsigned long long)ag << DS_RANK_BITS) | bg;
            unsigned long long f = (g << 1) | (nan ? 1ull : 0ull);
            if (f < d.first[S]) {
                unsigned long long old = atomicMin(&d.first[S], f);
                if (old == ~0ull) {
                    uint32_t pos = atomicAdd(&counters[k], 1u);
                    d.raw[pos] = S;
                }
            }
            if (!nan) {
                unsigned long long key = ((unsigned long long)ord_key(jr.total) << 32) |
                                         ((uint32_t)t << DS_RANK_BITS) | ag;
                if (key < d.best[S]) atomicMin(&d.best[S], key);
            }
        }
        __syncthreads();
    }
    add_pairs(counters, npairs);
    if (checked) atomicAdd((unsigned long long *)(counters + 44), (unsigned long long)checked);
}

__device__ __forceinline__ uint32_t ds_row(const DsBufs &d, unsigned long long f) {
    unsigned long long g = f >> 1;
    uint32_t t = (uint32_t)(g >> (2 * DS_RANK_BITS));
    uint32_t a = (uint32_t)(g >> DS_RANK_BITS) & ((1u << DS_RANK_BITS) - 1u);
    return d.sz[80 + t] + a;
}

__global__ void k_ds_rowcount(int k, DsBufs d, const uint32_t *counters) {
    uint32_t cnt = counters[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x)
        atomicAdd(&d.rowcnt[ds_row(d, d.first[d.raw[i]])], 1u);
}

/* exclusive scan of rowcnt[0..rows) into rowoff, one workgroup of 1024 */
__global__ void __launch_bounds__(1024) k_ds_scan(DsBufs d) {
    __shared__ uint32_t part[1024];
    const uint32_t rows = d.sz[160];
    const uint32_t per = (rows + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = b0; i < b0 + per && i < rows; i++) sum += d.rowcnt[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t acc = part[threadIdx.x] - sum;
    for (uint32_t i = b0; i < b0 + per && i < rows; i++) {
        d.rowoff[i] = acc;
        acc += d.rowcnt[i];
    }
}

__global__ void k_ds_scatter(int k, DsBufs d, const uint32_t *counters) {
    uint32_t cnt = counters[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
        uint32_t S = d.raw[i];
        uint32_t row = ds_row(d, d.first[S]);
        d.tmp[d.rowoff[row] + atomicAdd(&d.rowfill[row], 1u)] = S;
    }
}

/* rows with several new sets: order them by the b of their first pair */
__global__ void k_ds_rowsort(DsBufs d) {
    const uint32_t rows = d.sz[160];
    const unsigned long long bm = (1ull << DS_RANK_BITS) - 1ull;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        uint32_t c = d.rowcnt[r];
        if (c < 2) continue;
        uint32_t *seg = d.tmp + d.rowoff[r];
        for (uint32_t i = 1; i < c; i++) {
            uint32_t x = seg[i];
            unsigned long long bx = (d.first[x] >> 1) & bm;
            uint32_t j = i;
            while (j > 0 && ((d.first[seg[j - 1]] >> 1) & bm) > bx) {
                seg[j] = seg[j - 1];
                j--;
            }
            seg[j] = x;
        }
    }
}

__global__ void __launch_bounds__(BLOCK) k_ds_finalize(BlobHdr h, const uint32_t *blob, int k, DsBufs d,
                                                        float4 *vals, uint32_t *left, uint32_t *counters) {
    extern __shared__ GQ_LDS_ALIGN uint32_t smem[];
    for (int i = threadIdx.x; i < h.words; i += BLOCK) smem[i] = blob[i];
    __syncthreads();
    View<uint32_t> q = make_view(h, smem);
    const uint32_t *adj = smem + h.o_adj;
    const uint32_t cnt = counters[k];
    const uint32_t base = d.sz[40 + k];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < cnt; p += gridDim.x * blockDim.x) {
        uint32_t S = d.tmp[p];
        uint32_t nbr = 0, t = S;
        while (t) {
            int i = __builtin_ctz(t);
            t &= t - 1;
            nbr |= adj[i];
        }
        d.lset[base + p] = S;
        d.lnbr[base + p] = nbr;
        unsigned long long f = d.first[S];
        uint32_t tt, a;
        if (f & 1ull) {
            unsigned long long g = f >> 1;
            tt = (uint32_t)(g >> (2 * DS_RANK_BITS));
            a = (uint32_t)(g >> DS_RANK_BITS) & ((1u << DS_RANK_BITS) - 1u);
        } else {
            uint32_t r32 = (uint32_t)(d.best[S] & 0xffffffffull);
            tt = r32 >> DS_RANK_BITS;
            a = r32 & ((1u << DS_RANK_BITS) - 1u);
        }
        int ls = k - 1 - (int)tt;
        uint32_t l = d.lset[d.sz[40 + ls] + a], r = S ^ l;
        Rel jr = join(q, l, load_rel(vals, l), r, load_rel(vals, r));
        store_rel(vals, S, jr, q.p.work_mem);
        left[S] = l;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) d.sz[k] = cnt;
}

/* Pre-order walk of the plan from `root_set` (dpsub_buildQueryTree,
 * gpuqo_query_tree.cuh:46-72, done on the device in one pass). */
__global__ void __launch_bounds__(WAVE) k_tree(MemoRef mr, uint32_t root_set, DPNode *nodes, int cap,
                                               uint32_t *status) {
    const uint32_t r = walk_tree_wave(mr, root_set, nodes, cap);
    if (threadIdx.x == 0) status[0] = r;
}

/* IDP1 (gpuqo_idp.cu:38-80): the DP stops at level K and its result is the
 * cheapest set of K relations.  The reference scans its memo keeping the
 * first set whose cost total is strictly below the best so far, starting
 * from +inf (cpu_sequential.cu:101-112; gpuqo_dpsub_common.cu:350-358), so
 * NaN and +inf totals never win.  out[0] = the least order key of the
 * level-K totals below +inf (~0: none), out[1] = how many sets have it,
 * out[2 ..] = the first `cap` of them (the host orders ties).  The level's
 * list is list[*list_off ..] (list_off may be null) of *count sets. */
#define IDP1_TIES 64
__global__ void __launch_bounds__(1024) k_best_of_level(const uint32_t *list, const uint32_t *list_off,
                                                        const uint32_t *count, MemoRef mr, uint32_t *out,
                                                        uint32_t cap) {
    __shared__ uint32_t s_min, s_n;
    const uint32_t off = list_off ? *list_off : 0u, cnt = *count;
    const float inf = __builtin_huge_valf();
    if (threadIdx.x == 0) {
        s_min = ~0u;
        s_n = 0u;
    }
    __syncthreads();
    uint32_t best = ~0u;
    for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
        const float t = mr.vals[memo_slot(mr, list[off + i])].z;
        if (t < inf) {
            const uint32_t key = ord_key(t);
            best = key < best ? key : best;
        }
    }
    atomicMin(&s_min, best);
    __syncthreads();
    const uint32_t m = s_min;
    for (uint32_t i = threadIdx.x; i < cnt && m != ~0u; i += blockDim.x) {
        const uint32_t S = list[off + i];
        const float t = mr.vals[memo_slot(mr, S)].z;
        if (t < inf && ord_key(t) == m) {
            const uint32_t pos = atomicAdd(&s_n, 1u);
            if (pos < cap) out[2 + pos] = S;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = s_n;
    }
}

/* ------------------------------------------------------------ host side */

static uint64_t binom64(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return r;
}

/* Tasks of a tree job at level k with `iters` wave-iterations per task
 * (task_sets on the device). */
static uint64_t tree_tasks(uint64_t sets, int k, uint32_t iters) {
    const uint64_t spt = (uint64_t)(BLOCK / WAVE) * tree_spw(k) * iters;
    return (sets + spt - 1) / spt;
}

/* Launch-shape knobs.  The single-query and batch paths read g_tune under
 * the device mutex (g_ctx.mu); the step engine takes a copy under
 * g_tune_mu; tuning_set holds both. */
static Tuning g_tune;
static std::mutex g_tune_mu;

/* Ordered csg-cmp pairs of a tree query: 2(k-1) per connected set of k. */
static uint64_t tree_pairs(const std::vector<uint64_t> &counts, size_t kmax = 64) {
    uint64_t p = 0;
    for (size_t k = 2; k < counts.size() && k <= kmax; k++) p += counts[k] * 2ull * (k - 1);
    return p;
}


struct DeviceCtx {
    int device = -1;
    hipStream_t stream = nullptr;
    /* second stream of the batch (general-graph jobs beside the tree jobs;
     * default priority) and the per-level events that order the two */
    hipStream_t stream2 = nullptr;
    hipEvent_t lvl_ev[2][34] = {};
    hipEvent_t gate_ev[34] = {};        /* level completions of a paced DP (LevelGate) */
    /* the 64 device counters are known to be zero (a fused tree DP's walk
     * leaves them so): the next fused tree DP needs no clearing launch */
    bool counters_clean = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int wall_khz = 0;                   /* rate of the device clock wall_clock64 reads */
    DevBufs b{};
    size_t cap_memo = 0, cap_left = 0, cap_lvl = 0, cap_best = 0, cap_first = 0, cap_blob = 0, cap_conn = 0,
           cap_nodes = 0, cap_counters = 0;
    std::vector<hipEvent_t> prof_ev;
    DsBufs ds{};
    int *l1 = nullptr;
    uint32_t *lvl_off = nullptr;
    size_t cap_lvl_off = 0;
    /* batches: the packet (jobs, filter chunk prefix, blobs) staged from
     * coherent pinned memory by k_stage, results written back to it */
    char *ws = nullptr;
    uint8_t *bpkt = nullptr, *h_bpkt = nullptr;
    uint32_t *results = nullptr, *bprefix = nullptr, *h_bres = nullptr;
    size_t cap_ws = 0, cap_bpkt = 0, cap_h_bpkt = 0, cap_h_bres = 0, cap_results = 0, cap_bprefix = 0;
    /* big MPDP sets (both paths) and the single-query job descriptor */
    BigEnt *big_ents = nullptr;
    uint32_t *big_prefix = nullptr;
    BatchJob *sjob = nullptr;
    size_t cap_big = 0, cap_big_prefix = 0, cap_sjob = 0;
    /* IDP1: every tied set of the last level when the first pass kept too few */
    uint32_t *idp1_ties = nullptr;
    size_t cap_idp1_ties = 0;
    /* pinned host staging: query blob + level offsets in, counters + plan out */
    uint32_t *h_stage = nullptr, *h_res = nullptr;
    size_t cap_h_stage = 0;
    /* open-addressing memo of the single-query path */
    uint32_t *hkeys = nullptr, *hslot = nullptr;
    size_t cap_hkeys = 0, cap_hslot = 0;
    size_t cap_ds_best = 0, cap_ds_first = 0, cap_ds_lset = 0, cap_ds_lnbr = 0, cap_ds_raw = 0, cap_ds_tmp = 0,
           cap_ds_rc = 0, cap_ds_rf = 0, cap_ds_ro = 0, cap_ds_sz = 0, cap_l1 = 0;
    std::mutex mu;
};

static DeviceCtx g_ctx;

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, #x); \
            return GQ_ERR_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

template <class T>
static int ensure(T **p, size_t &cap, size_t need) {
    if (*p && cap >= need) return GQ_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
    if (hipMalloc((void **)p, bytes) != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc of %zu bytes failed", bytes);
        return GQ_ERR_NOMEM;
    }
    cap = std::max<size_t>(need, 1);
    return GQ_OK;
}

template <class T>
static int ensure_host(T **p, size_t &cap, size_t need);

static int ctx_init(DeviceCtx &c, int device) {
    if (c.device == device && c.stream) return GQ_OK;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c.stream2, hipStreamNonBlocking));
    for (int a = 0; a < 2; a++)
        for (int k = 0; k < 34; k++) HIPCHK(hipEventCreateWithFlags(&c.lvl_ev[a][k], hipEventDisableTiming));
    for (int k = 0; k < 34; k++) HIPCHK(hipEventCreateWithFlags(&c.gate_ev[k], hipEventDisableTiming));
    HIPCHK(hipEventCreate(&c.ev0));
    HIPCHK(hipEventCreate(&c.ev1));
    HIPCHK(hipDeviceGetAttribute(&c.wall_khz, hipDeviceAttributeWallClockRate, device));
    {
        uint32_t tab[33 * 33];
        for (int n = 0; n <= 32; n++)
            for (int k = 0; k <= 32; k++) tab[n * 33 + k] = (uint32_t)binom64(n, k);
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), tab, sizeof(tab)));
    }
    if (device >= 0 && device < GQ_MAX_DEVICES && !g_scratch[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));   /* this device's copy of the array */
        g_scratch[device] = (uintptr_t)p;
    }
    c.device = device;
    /* counters[0..63] and the plan nodes right after them: one read-back */
    int rc = ensure(&c.b.counters, c.cap_counters, RES_WORDS);
    if (rc) return rc;
    c.b.nodes = (DPNode *)(c.b.counters + 64);
    /* coherent: the plan walk writes the results straight into it */
    HIPCHK(hipHostMalloc((void **)&c.h_res, RES_WORDS * 4, hipHostMallocMapped | hipHostMallocCoherent));
    /* pinned staging at its working size, so the allocation-free path never
     * grows it */
    return ensure_host(&c.h_stage, c.cap_h_stage, na::STAGE_WORDS);
}

/* The status word of a fused tree DP's walk (h_res[48]) before the walk has
 * written it; the walk writes a node count (at most 63) or ~0u. */
#define RES_PENDING 0xfffffffdu

/* Wait for the walk's status word in pinned memory instead of for the
 * stream: the last launch writes the plan, the counters and then the status
 * word, a few microseconds before the stream reports the launch complete.
 * Later work on the stream is ordered after that launch anyway.  Every 4096
 * polls the stream is queried, so a failed launch (or one that ended with no
 * status) ends the wait with an error. */
static int wait_status(DeviceCtx &c, hipStream_t s) {
    for (uint64_t i = 0;; i++) {
        if (__atomic_load_n(&c.h_res[48], __ATOMIC_ACQUIRE) != RES_PENDING) break;
        if ((i & 4095u) == 4095u) {
            const hipError_t q = hipStreamQuery(s);
            if (q == hipErrorNotReady) continue;
            (void)hipGetLastError();
            if (q != hipSuccess) {
                set_error("HIP error %s while waiting for the plan", hipGetErrorString(q));
                return GQ_ERR_DEVICE;
            }
            if (__atomic_load_n(&c.h_res[48], __ATOMIC_ACQUIRE) != RES_PENDING) break;
            set_error("the plan walk left no status");
            return GQ_ERR_DEVICE;
        }
        __builtin_ia32_pause();
    }
    (void)hipGetLastError();                  /* the stream's not-ready answers are no error */
    return GQ_OK;
}

double host_now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/* CHECK_FOR_INTERRUPTS between the levels of one DP (the reference polls it
 * per level and per scatter: gpuqo_dpsub_common.cu:81,127,156,300;
 * gpuqo_dpsize.cu:451,486,536,554).  The host tests the interrupt flag
 * before it enqueues each level.  A paced DP -- one predicted to run for
 * milliseconds (paced_dp) -- also keeps at most two levels in flight:
 * before level k is enqueued the host waits for level k-2 to finish,
 * polling the completion event and the flag, so an interrupt raised while
 * the DP runs stops it within two levels.  Levels already enqueued always
 * run to completion (no kernel is cut short); the call then returns
 * GQ_ERR_INTERRUPTED.  The poll yields instead of sleeping: level k-1 is
 * still queued behind k-2, so the device stays busy while the host reacts. */
struct LevelGate {
    DeviceCtx &c;
    hipStream_t s;
    bool pace;
    int before(int k) {
        if (pace && k >= 4 && k - 2 < 34) {
            for (;;) {
                const hipError_t q = hipEventQuery(c.gate_ev[k - 2]);
                if (q == hipSuccess) break;
                if (q != hipErrorNotReady) {
                    set_error("HIP error %s while waiting for level %d", hipGetErrorString(q), k - 2);
                    return GQ_ERR_DEVICE;
                }
                if (interrupted()) break;
                std::this_thread::yield();
            }
        }
        if (interrupted()) {
            (void)hipStreamSynchronize(s);
            set_error("interrupted before level %d", k);
            return GQ_ERR_INTERRUPTED;
        }
        return GQ_OK;
    }
    int after(int k) {
        if (pace && k < 34) HIPCHK(hipEventRecord(c.gate_ev[k], s));
        return GQ_OK;
    }
};

/* DPs predicted to run for milliseconds, which are paced: tree DPs with
 * more than 10^8 csg-cmp pairs (exact, from the level sizes); general
 * graphs of at least 16 relations and at least 2n join edges (a 16-clique
 * plans in about 5 ms; sparse graphs such as cycles stay sub-millisecond);
 * DPsize from 14 relations with at least 2n edges, or 16 relations. */
static bool paced_dp(const DPInput &in, DPAlgo algo) {
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    uint64_t deg = 0;
    for (int v = 0; v < n; v++) deg += (uint64_t)__builtin_popcount(in.blob[(size_t)h.o_adj + v]);
    const bool dense = deg / 2 >= 2ull * (uint64_t)n;
    if (algo == DP_DPSIZE) return n >= 16 || (n >= 14 && dense);
    if (h.is_tree) {
        if ((int)in.tree_counts.size() != n + 1) return n >= 24;
        uint64_t pairs = 0;
        for (int k = 2; k <= n; k++) pairs += 2ull * (uint64_t)(k - 1) * in.tree_counts[k];
        return pairs > 100000000ull;
    }
    return n >= 16 && dense;
}

/* pinned staging buffer of at least `words` words */
/* Pinned host memory that kernels read and write directly: coherent, so
 * that no GPU cache keeps a line of it between two uses. */
template <class T>
static int ensure_host(T **p, size_t &cap, size_t need) {
    if (*p && cap >= need) return GQ_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    const size_t n = std::max<size_t>(std::max<size_t>(need, 16), 2 * cap);
    cap = 0;
    if (hipHostMalloc((void **)p, n * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        *p = nullptr;
        set_error("pinned allocation of %zu bytes failed", n * sizeof(T));
        return GQ_ERR_NOMEM;
    }
    cap = n;
    return GQ_OK;
}

static int ensure_stage(DeviceCtx &c, size_t words) {
    return ensure_host(&c.h_stage, c.cap_h_stage, words);
}

/* The staged words of the single-query path (query blob, level offsets, ...)
 * from c.h_stage to `dst` in one copy.  (A lone DP pays one copy-engine
 * hand-off at each end; measured, that is shorter than a staging kernel
 * reading host memory.) */
static int stage_up(DeviceCtx &c, hipStream_t s, uint32_t *dst, size_t words) {
    HIPCHK(hipMemcpyAsync(dst, c.h_stage, words * 4, hipMemcpyHostToDevice, s));
    return GQ_OK;
}

/* The DP's result (64 counters, then `nodes` plan nodes) into c.h_res. */
static int results_down(DeviceCtx &c, hipStream_t s, int nodes, const uint32_t *src = nullptr) {
    HIPCHK(hipMemcpyAsync(c.h_res, src ? src : c.b.counters, (64 + (size_t)nodes * (sizeof(DPNode) / 4)) * 4,
                          hipMemcpyDeviceToHost, s));
    return GQ_OK;
}

int device_init(int device) {
    std::lock_guard<std::mutex> g(g_ctx.mu);
    return ctx_init(g_ctx, device);
}

bool device_ready(int device) {
    std::lock_guard<std::mutex> g(g_ctx.mu);
    return g_ctx.device == device && g_ctx.stream != nullptr && device >= 0 && device < GQ_MAX_DEVICES &&
           g_scratch[device] != 0;
}

int with_device_stream(const std::function<int(int, void *)> &f) {
    std::lock_guard<std::mutex> g(g_ctx.mu);
    int rc = ctx_init(g_ctx, g_ctx.device < 0 ? 0 : g_ctx.device);
    if (rc) return rc;
    return f(g_ctx.device, (void *)g_ctx.stream);
}


/* ------------------------------------------------------------ batches */


bool fused_eligible(const DPInput &in, DPAlgo algo) {
    const BlobHdr &h = in.hdr;
    if (algo == DP_DPSIZE || h.n_rels < 2 || h.n_rels > 30) return false;
    if (algo == DP_DPSUB && !h.is_tree && h.n_rels > 12) return false;
    return std::max(eval_lds_bicc(h.words), eval_lds_dpsub(h.words, 0)) <= 64 * 1024;
}

static int run_dp_batch_impl(const std::vector<const DPInput *> &ins, const std::vector<DPAlgo> &algos,
                             std::vector<std::vector<DPNode>> &trees, std::vector<DPStats> &stats, double *gpu_ms) {
    DeviceCtx &c = g_ctx;
    int rc = ctx_init(c, c.device < 0 ? 0 : c.device);
    if (rc) return rc;
    const int nj = (int)ins.size();
    trees.assign(nj, std::vector<DPNode>());
    stats.assign(nj, DPStats());
    if (nj == 0) return GQ_OK;
    const Tuning tn = g_tune;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<uint32_t> chunk_prefix(nj + 1, 0);
    std::vector<size_t> ws_off(nj), blob_off(nj);
    size_t blob_words = 0;
    size_t ws = 0, lds = 0, lds_tree = 0, big_lds_max = 0;
    uint64_t big_cap = 0;
    int kmax = 2, big_kmax = 0;
    for (int j = 0; j < nj; j++) {
        const BlobHdr &h = ins[j]->hdr;
        if (!fused_eligible(*ins[j], algos[j])) {
            set_error("DP %d cannot run in a batch", j);
            return GQ_ERR_INVALID;
        }
        const uint64_t ns = 1ull << h.n_rels;
        ws_off[j] = ws;
        ws += al(ns * 16) + al(ns * 4) + al(std::max<uint64_t>(1, ns / 32) * 4) + al(ns * 4) + al(ns * 8) +
              al(ns * 4);
        blob_off[j] = blob_words;
        blob_words += (ins[j]->blob.size() + 63) & ~(size_t)63;
        if (fast_tree_job(h) && (int)ins[j]->tree_counts.size() != h.n_rels + 1) {
            set_error("tree job %d without level sizes", j);
            return GQ_ERR_INVALID;
        }
        if (fast_tree_job(h)) lds_tree = std::max(lds_tree, tree_lds(h.words, 1));
        else lds = std::max(lds, std::max(eval_lds_bicc(h.words), eval_lds_dpsub(h.words, 0)));
        if (algos[j] == DP_BICC && !h.is_tree && h.n_rels >= 15) {
            uint64_t mc = 0;
            for (int k = 2; k <= h.n_rels; k++) mc = std::max(mc, binom64(h.n_rels, k));
            big_cap += mc;
            big_lds_max = std::max(big_lds_max, big_lds(h.words));
            big_kmax = std::max(big_kmax, h.n_rels);
        }
        chunk_prefix[j + 1] = chunk_prefix[j] + (uint32_t)((ns + FILT_CHUNK - 1) / FILT_CHUNK);
        kmax = std::max(kmax, h.n_rels);
    }
    /* packet: jobs | chunk prefix | blobs (256-byte aligned parts) */
    const size_t o_pref = al(sizeof(BatchJob) * (size_t)nj), o_blob = al(o_pref + 4 * ((size_t)nj + 1));
    const size_t pkt_bytes = al(o_blob + 4 * blob_words);
    if ((rc = ensure(&c.ws, c.cap_ws, ws))) return rc;
    if ((rc = ensure(&c.bpkt, c.cap_bpkt, pkt_bytes))) return rc;
    if ((rc = ensure(&c.results, c.cap_results, (size_t)nj * RES_WORDS))) return rc;
    if ((rc = ensure_host(&c.h_bpkt, c.cap_h_bpkt, pkt_bytes))) return rc;
    if ((rc = ensure_host(&c.h_bres, c.cap_h_bres, (size_t)nj * RES_WORDS))) return rc;
    BatchJob *jobs = (BatchJob *)c.h_bpkt;
    memcpy(c.h_bpkt + o_pref, chunk_prefix.data(), 4 * ((size_t)nj + 1));
    for (int j = 0; j < nj; j++) {
        uint32_t *dst = (uint32_t *)(c.h_bpkt + o_blob) + blob_off[j];
        const std::vector<uint32_t> &b = ins[j]->blob;
        memcpy(dst, b.data(), 4 * b.size());
        memset(dst + b.size(), 0, 4 * (((b.size() + 63) & ~(size_t)63) - b.size()));
    }
    BatchJob *const d_jobs = (BatchJob *)c.bpkt;
    if ((rc = ensure(&c.bprefix, c.cap_bprefix, (size_t)(nj + 1) * 69))) return rc;
    if (big_cap > 0xffffffffull) {
        set_error("batch too large for the big-set list");
        return GQ_ERR_UNSUPPORTED;
    }
    if (big_cap) {
        if ((rc = ensure(&c.big_ents, c.cap_big, big_cap))) return rc;
        if ((rc = ensure(&c.big_prefix, c.cap_big_prefix, big_cap + 1))) return rc;
    }
    for (int j = 0; j < nj; j++) {
        const BlobHdr &h = ins[j]->hdr;
        const uint64_t ns = 1ull << h.n_rels;
        BatchJob &J = jobs[j];
        memset(&J, 0, sizeof(J));
        J.h = h;
        J.blob = (const uint32_t *)(c.bpkt + o_blob) + blob_off[j];
        char *p = c.ws + ws_off[j];
        J.vals = (float4 *)p;
        p += al(ns * 16);
        J.left = (uint32_t *)p;
        p += al(ns * 4);
        J.conn = (uint32_t *)p;
        J.conn_words = (uint32_t)std::max<uint64_t>(1, ns / 32);
        p += al(std::max<uint64_t>(1, ns / 32) * 4);
        J.lvl_sets = (uint32_t *)p;
        p += al(ns * 4);
        J.lvl_best = (unsigned long long *)p;
        p += al(ns * 8);
        J.lvl_first = (uint32_t *)p;
        J.counters = c.results + (size_t)j * RES_WORDS;
        J.nodes = (DPNode *)(J.counters + 64);
        uint64_t acc = 0;
        for (int k = 0; k <= 32; k++) {
            J.lvl_off[k] = (uint32_t)acc;
            if (k >= 2 && k <= h.n_rels) acc += binom64(h.n_rels, k);
        }
        J.bicc = dp_mode(*ins[j], algos[j]);
        J.tree_iters = tn.tree_iters;
    }
    hipStream_t s = c.stream;
    {
        const uint32_t n16 = (uint32_t)(pkt_bytes / 16);
        hipLaunchKernelGGL(k_stage, dim3(std::max<uint32_t>(1, std::min<uint32_t>((n16 + BLOCK - 1) / BLOCK, 512))),
                           dim3(BLOCK), 0, s, (const uint4 *)c.h_bpkt, (uint4 *)c.bpkt, n16);
    }
    HIPCHK(hipEventRecord(c.ev0, s));
    hipLaunchKernelGGL(k_batch_init, dim3(std::min(nj, 1024)), dim3(64), 0, s, d_jobs, nj, nullptr, nullptr);
    const uint32_t fgrid = std::min<uint32_t>(chunk_prefix[nj], 4096);
    hipLaunchKernelGGL(k_batch_filter, dim3(std::max<uint32_t>(fgrid, 1)), dim3(BLOCK), 0, s, d_jobs, nj,
                       (const uint32_t *)(c.bpkt + o_pref));
    /* task prefixes [33][nj+1] after the chunk prefix: tree jobs, then the
     * others.  Tree jobs run in k_batch_eval_tree on stream s, the others in
     * k_batch_eval on stream2; level k on either waits for level k-1 on both. */
    uint32_t *tp_tree = c.bprefix + (nj + 1), *tp_gen = c.bprefix + (size_t)(nj + 1) * 35;
    hipLaunchKernelGGL(k_batch_plan, dim3(kmax - 1), dim3(1024), 0, s, d_jobs, nj, tp_tree, JOBS_TREE);
    hipLaunchKernelGGL(k_batch_plan, dim3(kmax - 1), dim3(1024), 0, s, d_jobs, nj, tp_gen, JOBS_GENERAL);
    const size_t lds_tree_total = std::max<size_t>(lds_tree, 16);
    const uint32_t tpl_gen = (lds + 4 * (size_t)(nj + 1) <= 64 * 1024) ? (uint32_t)(lds / 4) : 0u;
    const size_t lds_gen_total = tpl_gen ? lds + 4 * (size_t)(nj + 1) : lds;
    uint32_t *big_hdr = c.results + 50;                  /* job 0's counters[50..51] */
    hipStream_t s2 = c.stream2;
    HIPCHK(hipEventRecord(c.lvl_ev[0][1], s));
    HIPCHK(hipStreamWaitEvent(s2, c.lvl_ev[0][1], 0));
    for (int k = 2; k <= kmax; k++) {
        if (interrupted()) {                       /* CHECK_FOR_INTERRUPTS per level */
            (void)hipStreamSynchronize(s);
            (void)hipStreamSynchronize(s2);
            set_error("interrupted before batch level %d", k);
            return GQ_ERR_INTERRUPTED;
        }
        /* grids: enough workgroups for the largest level, capped */
        uint64_t tasks_tree = 0, ub_gen = 0;
        for (int j = 0; j < nj; j++) {
            const BlobHdr &h = ins[j]->hdr;
            if (k > h.n_rels) continue;
            if (fast_tree_job(h)) tasks_tree += tree_tasks(ins[j]->tree_counts[k], k, tn.tree_iters);
            else {                        /* tasks of a bound on the level size */
                const uint64_t lvl = (h.is_tree && (int)ins[j]->tree_counts.size() == h.n_rels + 1)
                                         ? ins[j]->tree_counts[k] : binom64(h.n_rels, k);
                const uint64_t spt = task_sets_of(h.is_tree != 0, dp_mode(*ins[j], algos[j]) != 0, k, tn.tree_iters);
                ub_gen += (lvl + spt - 1) / spt;
            }
        }
        if (tasks_tree) {
            if (tasks_tree > 0x7fffffffull) {
                set_error("too many tasks in one level");
                return GQ_ERR_UNSUPPORTED;
            }
            hipLaunchKernelGGL(k_batch_eval_tree, dim3((uint32_t)tasks_tree), dim3(BLOCK), lds_tree_total, s, d_jobs,
                               nj, k, (const uint32_t *)(tp_tree + (size_t)k * (nj + 1)));
        }
        if (ub_gen) {
            /* its 192 VGPRs allow two workgroups per CU, 512 on the GPU; more
             * only queue behind the tree kernel running beside it, and the
             * launch ends when the last has been dispatched */
            const uint32_t grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(ub_gen, 1), tn.gen_grid_max);
            hipLaunchKernelGGL(k_batch_eval, dim3(grid), dim3(BLOCK), lds_gen_total, s2, d_jobs, nj, k,
                               tp_gen + (size_t)k * (nj + 1), tpl_gen, big_cap ? c.big_ents : nullptr,
                               (uint32_t)big_cap, big_hdr);
            if (big_cap && k >= 15 && k <= big_kmax) {
                hipLaunchKernelGGL(k_big_scan, dim3(1), dim3(1024), 0, s2, c.big_ents, big_hdr, (uint32_t)big_cap,
                                   c.big_prefix);
                hipLaunchKernelGGL(k_big_eval, dim3(1024), dim3(BLOCK), big_lds_max, s2, d_jobs, c.big_ents,
                                   big_hdr, c.big_prefix);
            }
        }
        HIPCHK(hipEventRecord(c.lvl_ev[0][k], s));
        HIPCHK(hipEventRecord(c.lvl_ev[1][k], s2));
        HIPCHK(hipStreamWaitEvent(s, c.lvl_ev[1][k], 0));
        HIPCHK(hipStreamWaitEvent(s2, c.lvl_ev[0][k], 0));
    }
    hipLaunchKernelGGL(k_batch_tree, dim3(nj), dim3(WAVE), 0, s, d_jobs, nj, nullptr, c.h_bres);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev1, s));
    HIPCHK(hipStreamSynchronize(s));
    const uint32_t *const res_h = c.h_bres;
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    if (gpu_ms) *gpu_ms = ms;
    for (int j = 0; j < nj; j++) {
        const int n = ins[j]->hdr.n_rels;
        const uint32_t *cnt = res_h + (size_t)j * RES_WORDS;
        if (cnt[49]) {
            set_error("a connected set has more than 2^31 splits (unsupported on the GPU path)");
            return GQ_ERR_UNSUPPORTED;
        }
        if (cnt[48] != (uint32_t)(2 * n - 1)) {
            set_error("plan reconstruction failed in batch job %d (status %u)", j, cnt[48]);
            return GQ_ERR_DEVICE;
        }
        const DPNode *nodes = (const DPNode *)(cnt + 64);
        trees[j].assign(nodes, nodes + 2 * n - 1);
        DPStats &st = stats[j];
        st.candidates = (1ull << n) - (uint64_t)n - 1ull;
        for (int k = 2; k <= n; k++) st.sets += cnt[k];
        st.pairs = (uint64_t)cnt[40] | ((uint64_t)cnt[41] << 32);
        if (ins[j]->hdr.is_tree && (int)ins[j]->tree_counts.size() == n + 1) st.pairs = tree_pairs(ins[j]->tree_counts);
        st.memo_bytes = (1ull << n) * 40;
        st.memo_kind = 1;
        st.gpu_ms = ms / nj;
        st.eval_ms = ms / nj;
        st.eval_launches = 1;
    }
    return GQ_OK;
}

int run_dp_batch(const std::vector<const DPInput *> &ins, const std::vector<DPAlgo> &algos,
                 std::vector<std::vector<DPNode>> &trees, std::vector<DPStats> &stats, double *gpu_ms) {
    std::lock_guard<std::mutex> g(g_ctx.mu);
    return run_dp_batch_impl(ins, algos, trees, stats, gpu_ms);
}



/* One level of a tree query of `sets` sets (an upper bound): the
 * per-edge-table kernel when the blob has the table, else the class-mask
 * kernel for its mask width; two lanes per edge when the level is small.
 * The grid covers the level, capped at 8192 workgroups. */
static void launch_eval_tree(const BlobHdr &h, uint64_t sets_k, size_t lds, hipStream_t s, const uint32_t *blob,
                             int k, int bicc, const uint32_t *sets, const uint32_t *slots, MemoRef mr,
                             uint32_t *counters, uint32_t *next = nullptr, uint32_t *next_slot = nullptr,
                             uint32_t flags = 0u, uint32_t *res_host = nullptr) {
    const uint64_t spw2 = tree_spw(k, 2);
    /* levels of at most g_tune.lpe2_waves waves at two lanes per edge leave
     * the GPU latency-bound (a few waves per SIMD): the two-lane form */
    const int lpe = (flags || (spw2 > 0 && (sets_k + spw2 - 1) / spw2 <= g_tune.lpe2_waves)) ? 2 : 1;
    const uint64_t spw = tree_spw(k, lpe);
    const uint64_t waves = (sets_k + spw - 1) / spw;
    /* two-lane form: one wave-iteration per wave (latency).  One-lane form:
     * each wave queues its winners' records and writes them, hash table
     * size packed, a wave at a time (the packing costs the same for 2 or 64
     * records), so each wave takes about g_tune.tree_iters iterations, on
     * at least 1024 workgroups while the level has them */
    uint64_t evw = (waves + 3) / 4;
    if (lpe == 1) {
        const uint64_t it = std::max<uint32_t>(1, g_tune.tree_iters);
        evw = std::max<uint64_t>(std::min<uint64_t>(evw, 1024), (waves + 4 * it - 1) / (4 * it));
    }
    const uint32_t ev = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(evw, 8192));
    /* with `next`: the level k+1 list from the sets_k (exact) level-k sets */
    const uint32_t ex = next ? (uint32_t)((sets_k + BLOCK - 1) / BLOCK) : 0u;
    const uint32_t xs = (g_tune.one_xcd_wgs && ev + ex <= g_tune.one_xcd_wgs) ? XCDS : 1u;
    const dim3 grid((ev + ex) * xs);
    lds = std::max(lds, tree_lds(h.words, lpe));
    if (flags & TREE_WALK) lds = std::max(lds, tree_lds(h.words, lpe) + (WALK_WS_WORDS + 8) * 4);
    const uint32_t cap = (uint32_t)std::min<uint64_t>(sets_k, 0xffffffffull);
#define GQ_TREE_LAUNCH(E, F, P, M)                                                                            \
    hipLaunchKernelGGL((k_eval_tree<E, F, P, M>), grid, dim3(BLOCK), lds, s, h, blob, k, bicc, sets, slots, mr, \
                       counters, ev, next, cap, next_slot, xs, res_host)
    /* the start and walk forms are kernels of their own (the common level
     * kernel keeps its smaller code); both run one-set-per-few-lanes levels
     * (level 2: n - 1 sets; the last level: one set), i.e. two lanes per edge */
#define GQ_TREE_MODES(E, F)                                                                                   \
    if (flags == (TREE_START | TREE_WALK)) GQ_TREE_LAUNCH(E, F, 2, TREE_START | TREE_WALK);                   \
    else if (flags == TREE_START) GQ_TREE_LAUNCH(E, F, 2, TREE_START);                                        \
    else if (flags == TREE_WALK) GQ_TREE_LAUNCH(E, F, 2, TREE_WALK);                                          \
    else if (lpe == 2) GQ_TREE_LAUNCH(E, F, 2, 0u);                                                           \
    else GQ_TREE_LAUNCH(E, F, 1, 0u)
    if (h.tree_fast) {
        GQ_TREE_MODES(1, true);
    } else if (h.ecw == 1) {
        GQ_TREE_MODES(1, false);
    } else {
        GQ_TREE_MODES(0, false);
    }
#undef GQ_TREE_MODES
#undef GQ_TREE_LAUNCH
}

static MemoRef dense_memo_of(const DevBufs &b) {
    MemoRef m;
    m.vals = b.vals;
    m.left = b.left;
    m.keys = nullptr;
    m.mask = 0;
    m.err = b.counters + 49;
    m.rank = nullptr;
    m.rn = 0;
    return m;
}
static MemoRef dense_memo(DeviceCtx &c) { return dense_memo_of(c.b); }

/* The allocation-free buffers of `device` (inside `scratch`). */
static int scratch_bufs(int device, DevBufs &b, BigEnt **big, uint32_t **big_prefix) {
    if (device < 0 || device >= GQ_MAX_DEVICES || !g_scratch[device]) {
        set_error("no scratch address for device %d (gq_init must run first)", device);
        return GQ_ERR_INVALID;
    }
    char *base = (char *)g_scratch[device];
    b.vals = (float4 *)(base + na::O_VALS);
    b.left = (uint32_t *)(base + na::O_LEFT);
    b.conn = (uint32_t *)(base + na::O_CONN);
    b.lvl_sets = (uint32_t *)(base + na::O_SETS);
    b.lvl_best = (unsigned long long *)(base + na::O_BEST);
    b.lvl_first = (uint32_t *)(base + na::O_FIRST);
    b.counters = (uint32_t *)(base + na::O_RES);
    b.nodes = (DPNode *)(b.counters + 64);
    b.blob = (uint32_t *)(base + na::O_BLOB);
    *big = (BigEnt *)(base + na::O_BIG);
    *big_prefix = (uint32_t *)(base + na::O_BIGP);
    return GQ_OK;
}

/* IDP1 ties: the first of `cands` in the iteration order of the CPU DP's
 * memo -- a libstdc++ unordered_map keyed by the reference's relids
 * (relation i is bit i+1) with the identity hash, filled with the base
 * relations in index order and then with every set of 2..K relations in
 * the order the DP creates them, `created` (cpu_sequential.cu:27-129).
 * The same container, keys and insertion sequence give the same order. */
static uint32_t memo_order_first(int n, const std::vector<uint32_t> &created,
                                 const std::unordered_set<uint32_t> &cands) {
    std::unordered_map<uint64_t, char> m;
    for (int i = 0; i < n; i++) m.emplace(1ull << (i + 1), 0);
    for (uint32_t S : created) m.emplace((uint64_t)S << 1, 0);
    for (const auto &kv : m)
        if (cands.count((uint32_t)(kv.first >> 1))) return (uint32_t)(kv.first >> 1);
    return 0u;
}

/* The sets of 2..K relations in cpu_dpsub's / cpu_dpsub_bicc's creation
 * order under IDP: level by level, connected sets in increasing integer
 * order (gpuqo_cpu_dpsub.cuh:26-63). */
static std::vector<uint32_t> dpsub_creation_order(const DPInput &in, int K) {
    const int n = in.hdr.n_rels;
    const uint32_t *adj = in.blob.data() + in.hdr.o_adj;
    std::vector<uint32_t> out;
    for (int k = 2; k <= K; k++) {
        uint64_t S = (1ull << k) - 1ull;
        const uint64_t last = S << (n - k);
        while (true) {
            const uint32_t s32 = (uint32_t)S;
            uint32_t seen = s32 & (0u - s32), front = seen;
            while (front) {
                uint32_t nb = 0;
                for (uint32_t f = front; f; f &= f - 1u) nb |= adj[__builtin_ctz(f)];
                front = nb & s32 & ~seen;
                seen |= front;
            }
            if (seen == s32) out.push_back(s32);
            if (S == last) break;
            const uint64_t c = S & (0ull - S), r = S + c;             /* next k-subset (Gosper) */
            S = (((r ^ S) >> 2) / c) | r;
        }
    }
    return out;
}

/* Largest IDP1 memo (sets of 1..K relations, counted by C(n, k)) the host
 * replays to order tied sets; above it a tie is broken by the lowest set,
 * which may differ from the CPU DP's choice (its memo iteration order). */
#define IDP1_REPLAY_MAX (1ull << 24)

/* IDP1: the plan root, the cheapest set of level K (k_best_of_level) with
 * ties ordered as the CPU DP orders them.  `created` produces the creation
 * order when there is a tie. */
template <class Created>
static int idp1_root(DeviceCtx &c, hipStream_t s, const MemoRef &mr, const uint32_t *list, const uint32_t *list_off,
                     const uint32_t *count, int n, int K, Created created, uint32_t *root) {
    uint32_t *dout = (uint32_t *)c.b.nodes;                  /* free until the plan walk */
    hipLaunchKernelGGL(k_best_of_level, dim3(1), dim3(1024), 0, s, list, list_off, count, mr, dout,
                       (uint32_t)IDP1_TIES);
    HIPCHK(hipGetLastError());
    uint32_t *r = c.h_res + 64;
    HIPCHK(hipMemcpyAsync(r, dout, (2 + IDP1_TIES) * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (r[0] == ~0u || r[1] == 0) {
        set_error("IDP1: no set of %d relations with a cost below +inf", K);
        return GQ_ERR_NO_PLAN;
    }
    if (r[1] == 1) {
        *root = r[2];
        return GQ_OK;
    }
    std::vector<uint32_t> ties(r + 2, r + 2 + std::min<uint32_t>(r[1], IDP1_TIES));
    if (r[1] > IDP1_TIES) {                                  /* more ties than the first pass kept: all of them */
        const uint32_t nt = r[1];
        int rc0 = ensure(&c.idp1_ties, c.cap_idp1_ties, (size_t)nt + 2);
        if (rc0) return rc0;
        hipLaunchKernelGGL(k_best_of_level, dim3(1), dim3(1024), 0, s, list, list_off, count, mr, c.idp1_ties, nt);
        HIPCHK(hipGetLastError());
        ties.resize((size_t)nt + 2);
        HIPCHK(hipMemcpyAsync(ties.data(), c.idp1_ties, ((size_t)nt + 2) * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (ties[1] != nt) {
            set_error("IDP1: the tie count changed between passes");
            return GQ_ERR_DEVICE;
        }
        ties.erase(ties.begin(), ties.begin() + 2);
    }
    uint64_t sets = 0;
    for (int k = 1; k <= K; k++) sets += binom64(n, k);
    if (sets > IDP1_REPLAY_MAX) {
        *root = *std::min_element(ties.begin(), ties.end());
        return GQ_OK;
    }
    const std::unordered_set<uint32_t> cands(ties.begin(), ties.end());
    std::vector<uint32_t> order;
    int rc = created(order);
    if (rc) return rc;
    *root = memo_order_first(n, order, cands);
    if (!*root) {
        set_error("IDP1: tied sets missing from the replayed memo");
        return GQ_ERR_DEVICE;
    }
    return GQ_OK;
}

/* Tree graphs with the open-addressing memo: the level lists are generated
 * by expansion (k_expand) instead of filtering all 2^n integers, and the
 * memo holds only the connected sets (capacity: the next power of two of
 * twice their number, so probes stay short). */
static int run_dp_hash(DeviceCtx &c, const DPInput &in, DPAlgo algo, int K, bool profile,
                       std::vector<DPNode> &tree, DPStats &st) {
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    int rc;
    if (!h.is_tree || (int)in.tree_counts.size() != n + 1 || algo == DP_DPSIZE) {
        set_error("the open-addressing memo serves DPsub/MPDP on tree graphs");
        return GQ_ERR_UNSUPPORTED;
    }
    uint64_t total = 0, list = 0;
    uint32_t off_h[33], tasks[33];
    memset(off_h, 0, sizeof(off_h));
    memset(tasks, 0, sizeof(tasks));
    for (int k = 1; k <= n; k++) {
        total += in.tree_counts[k];
        if (k >= 2) {
            off_h[k] = (uint32_t)list;
            list += in.tree_counts[k];
            tasks[k] = (uint32_t)((in.tree_counts[k - 1] + BLOCK - 1) / BLOCK);
        }
    }
    if (2 * total > (1ull << 31)) {
        set_error("too many connected sets for the open-addressing memo (%llu)", (unsigned long long)total);
        return GQ_ERR_UNSUPPORTED;
    }
    uint64_t cap = 64;
    while (cap < 2 * total) cap <<= 1;
    const size_t W4h = ((size_t)h.words + 3) & ~(size_t)3;
    const size_t lds_eval = W4h * 4;
    if ((rc = ensure(&c.hkeys, c.cap_hkeys, cap))) return rc;
    if ((rc = ensure(&c.b.vals, c.cap_memo, cap))) return rc;
    if ((rc = ensure(&c.b.left, c.cap_left, cap))) return rc;
    if ((rc = ensure(&c.b.lvl_sets, c.cap_lvl, std::max<uint64_t>(list, 1)))) return rc;
    if ((rc = ensure(&c.hslot, c.cap_hslot, std::max<uint64_t>(list, 1)))) return rc;
    /* staged area: blob | level offsets (33, padded to 36) | task prefixes
     * (66, padded to 68) | the job descriptor */
    const size_t o_pre = W4h + 36, o_sj = o_pre + 68;
    const size_t stage_words = o_sj + ((sizeof(BatchJob) + 15) & ~(size_t)15) / 4;
    if ((rc = ensure(&c.b.blob, c.cap_blob, stage_words))) return rc;
    if ((rc = ensure_stage(c, stage_words))) return rc;
    c.lvl_off = c.b.blob + W4h;
    uint32_t *d_prefix = c.b.blob + o_pre;              /* [k] -> {0, tasks[k]} */
    BatchJob *const d_sj = (BatchJob *)(c.b.blob + o_sj);
    hipStream_t s = c.stream;
    memset(c.h_stage, 0, stage_words * 4);
    memcpy(c.h_stage, in.blob.data(), (size_t)h.words * 4);
    memcpy(c.h_stage + W4h, off_h, sizeof(off_h));
    for (int k = 0; k < 33; k++) {
        c.h_stage[o_pre + 2 * k] = 0;
        c.h_stage[o_pre + 2 * k + 1] = tasks[k];
    }
    BatchJob &sj = *(BatchJob *)(c.h_stage + o_sj);
    sj.h = h;
    sj.blob = c.b.blob;
    sj.vals = c.b.vals;
    sj.left = c.b.left;
    sj.lvl_sets = c.b.lvl_sets;
    sj.lvl_slot = c.hslot;
    sj.keys = c.hkeys;
    sj.hmask = (uint32_t)(cap - 1);
    sj.counters = c.b.counters;
    sj.nodes = c.b.nodes;
    memcpy(sj.lvl_off, off_h, sizeof(off_h));
    sj.bicc = algo == DP_BICC ? 1 : 0;
    st.t_first_launch = host_now_ms();
    if ((rc = stage_up(c, s, c.b.blob, stage_words))) return rc;
    MemoRef mr;
    mr.vals = c.b.vals;
    mr.left = c.b.left;
    mr.keys = c.hkeys;
    mr.mask = (uint32_t)(cap - 1);
    mr.err = c.b.counters + 49;
    mr.rank = nullptr;
    mr.rn = 0;
    if (profile && (int)c.prof_ev.size() < 3 * 33) {
        while ((int)c.prof_ev.size() < 3 * 33) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c.prof_ev.push_back(e);
        }
    }
    HIPCHK(hipEventRecord(c.ev0, s));
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)std::min<uint64_t>((cap + BLOCK - 1) / BLOCK, 4096)), dim3(BLOCK),
                       0, s, c.hkeys, cap, 0u);
    hipLaunchKernelGGL(k_batch_init, dim3(1), dim3(64), 0, s, d_sj, 1, nullptr, nullptr);
    LevelGate gate{c, s, paced_dp(in, algo)};
    for (int k = 2; k <= K; k++) {
        if ((rc = gate.before(k))) return rc;
        if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k], s));
        hipLaunchKernelGGL(k_expand, dim3(std::max<uint32_t>(1, std::min<uint32_t>(tasks[k], 4096))), dim3(BLOCK), 0,
                           s, d_sj, 1, k, d_prefix + 2 * k);
        if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 1], s));
        launch_eval_tree(h, in.tree_counts[k], lds_eval, s, c.b.blob, k, sj.bicc, c.b.lvl_sets + off_h[k],
                         c.hslot + off_h[k], mr, c.b.counters);
        if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 2], s));
        if ((rc = gate.after(k))) return rc;
    }
    uint32_t root = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
    if (K < n) {                                   /* IDP1: the cheapest set of K relations */
        rc = idp1_root(c, s, mr, c.b.lvl_sets + off_h[K], nullptr, c.b.counters + K, n, K,
                       [&](std::vector<uint32_t> &o) -> int { o = dpsub_creation_order(in, K); return GQ_OK; }, &root);
        if (rc) return rc;
    }
    const int nodes_n = 2 * K - 1;
    hipLaunchKernelGGL(k_tree, dim3(1), dim3(64), 0, s, mr, root, c.b.nodes, nodes_n, c.b.counters + 48);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev1, s));
    tree.resize(nodes_n);
    const uint32_t *cnt = c.h_res;
    st.t_last_launch = host_now_ms();
    if ((rc = results_down(c, s, nodes_n))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    st.t_synced = host_now_ms();
    if (cnt[49]) {
        set_error("open-addressing memo error (flags %u)", cnt[49]);
        return (cnt[49] & 1u) ? GQ_ERR_UNSUPPORTED : GQ_ERR_DEVICE;
    }
    for (int k = 2; k <= K; k++)
        if (cnt[k] != in.tree_counts[k]) {
            set_error("level %d holds %u sets, expected %llu", k, cnt[k], (unsigned long long)in.tree_counts[k]);
            return GQ_ERR_DEVICE;
        }
    if (cnt[48] != (uint32_t)nodes_n) {
        set_error("plan reconstruction failed (memo inconsistent, status %u)", cnt[48]);
        return GQ_ERR_DEVICE;
    }
    memcpy(tree.data(), c.h_res + 64, sizeof(DPNode) * nodes_n);
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    st.gpu_ms = ms;
    if (profile) {
        for (int k = 2; k <= K; k++) {
            float a = 0.0f, b = 0.0f;
            HIPCHK(hipEventElapsedTime(&a, c.prof_ev[3 * k], c.prof_ev[3 * k + 1]));
            HIPCHK(hipEventElapsedTime(&b, c.prof_ev[3 * k + 1], c.prof_ev[3 * k + 2]));
            st.filter_ms += a;
            st.eval_ms += b;
            st.filter_launches++;
            st.eval_launches++;
        }
    }
    uint64_t made = 0;
    for (int k = 2; k <= K; k++) made += in.tree_counts[k];
    st.candidates = made;            /* sets generated by expansion (no 2^n scan) */
    st.sets = made;
    st.pairs = tree_pairs(in.tree_counts, (size_t)K);
    st.memo_bytes = cap * (sizeof(float4) + 4 + 4) + list * 8;
    st.memo_kind = GQ_MEMO_HASH;
    return GQ_OK;
}

/* Tree graphs with the rank memo: one slot per connected set, computed from
 * the set (rank_slot), so the memo holds 20 bytes per connected set and no
 * keys.  Level 2 comes from k_tree_start and each level's launch builds the
 * next level's list by expansion, as in the dense-memo tree DP; the
 * expansion also writes each new set's slot (its parent's slot plus one
 * table entry), which the next level's evaluation reads with the set. */
static int run_dp_rank(DeviceCtx &c, const DPInput &in, DPAlgo algo, int K, bool profile,
                       std::vector<DPNode> &tree, DPStats &st) {
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    int rc;
    if (!h.is_tree || !h.o_rank || (int)in.tree_counts.size() != n + 1 || algo == DP_DPSIZE || n < 2) {
        set_error("the rank memo serves DPsub/MPDP on tree graphs");
        return GQ_ERR_UNSUPPORTED;
    }
    uint64_t total = 0, list = 0;
    uint32_t off_h[33];
    memset(off_h, 0, sizeof(off_h));
    for (int k = 1; k <= n; k++) {
        total += in.tree_counts[k];
        if (k >= 2) {
            off_h[k] = (uint32_t)list;
            list += in.tree_counts[k];
        }
    }
    if (total >= (1ull << 31)) {
        set_error("too many connected sets for the rank memo (%llu)", (unsigned long long)total);
        return GQ_ERR_UNSUPPORTED;
    }
    const size_t W4h = ((size_t)h.words + 3) & ~(size_t)3;
    const size_t lds_eval = W4h * 4;
    if ((rc = ensure(&c.b.vals, c.cap_memo, total))) return rc;
    if ((rc = ensure(&c.b.left, c.cap_left, total))) return rc;
    if ((rc = ensure(&c.b.lvl_sets, c.cap_lvl, std::max<uint64_t>(list, 1)))) return rc;
    if ((rc = ensure(&c.hslot, c.cap_hslot, std::max<uint64_t>(list, 1)))) return rc;
    if ((rc = ensure(&c.b.blob, c.cap_blob, W4h + 36))) return rc;
    if ((rc = ensure_stage(c, W4h + 36))) return rc;
    c.lvl_off = c.b.blob + W4h;
    hipStream_t s = c.stream;
    memset(c.h_stage, 0, (W4h + 36) * 4);
    memcpy(c.h_stage, in.blob.data(), (size_t)h.words * 4);
    memcpy(c.h_stage + W4h, off_h, sizeof(off_h));
    st.t_first_launch = host_now_ms();
    if ((rc = stage_up(c, s, c.b.blob, W4h + 36))) return rc;
    MemoRef mr = dense_memo(c);
    mr.rank = c.b.blob + h.o_rank;
    mr.rn = n;
    if (profile && (int)c.prof_ev.size() < 3 * 33) {
        while ((int)c.prof_ev.size() < 3 * 33) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c.prof_ev.push_back(e);
        }
    }
    const int bicc = algo == DP_BICC ? 1 : 0;
    HIPCHK(hipEventRecord(c.ev0, s));
    hipLaunchKernelGGL(k_tree_start, dim3(1), dim3(WAVE), 0, s, h, c.b.blob, c.b.vals, c.b.left,
                       c.b.lvl_sets + off_h[2], c.b.counters, c.hslot + off_h[2]);
    LevelGate gate{c, s, paced_dp(in, algo)};
    for (int k = 2; k <= K; k++) {
        if ((rc = gate.before(k))) return rc;
        if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 1], s));
        launch_eval_tree(h, in.tree_counts[k], lds_eval, s, c.b.blob, k, bicc, c.b.lvl_sets + off_h[k],
                         c.hslot + off_h[k], mr, c.b.counters, k < K ? c.b.lvl_sets + off_h[k + 1] : nullptr,
                         k < K ? c.hslot + off_h[k + 1] : nullptr);
        if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 2], s));
        if ((rc = gate.after(k))) return rc;
    }
    uint32_t root = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
    if (K < n) {                                   /* IDP1: the cheapest set of K relations */
        rc = idp1_root(c, s, mr, c.b.lvl_sets + off_h[K], nullptr, c.b.counters + K, n, K,
                       [&](std::vector<uint32_t> &o) -> int { o = dpsub_creation_order(in, K); return GQ_OK; }, &root);
        if (rc) return rc;
    }
    const int nodes_n = 2 * K - 1;
    hipLaunchKernelGGL(k_tree, dim3(1), dim3(64), 0, s, mr, root, c.b.nodes, nodes_n, c.b.counters + 48);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev1, s));
    tree.resize(nodes_n);
    const uint32_t *cnt = c.h_res;
    st.t_last_launch = host_now_ms();
    if ((rc = results_down(c, s, nodes_n))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    st.t_synced = host_now_ms();
    if (cnt[49]) {
        set_error("rank memo error (flags %u)", cnt[49]);
        return (cnt[49] & 1u) ? GQ_ERR_UNSUPPORTED : GQ_ERR_DEVICE;
    }
    for (int k = 2; k <= K; k++)
        if (cnt[k] != in.tree_counts[k]) {
            set_error("level %d holds %u sets, expected %llu", k, cnt[k], (unsigned long long)in.tree_counts[k]);
            return GQ_ERR_DEVICE;
        }
    if (cnt[48] != (uint32_t)nodes_n) {
        set_error("plan reconstruction failed (memo inconsistent, status %u)", cnt[48]);
        return GQ_ERR_DEVICE;
    }
    memcpy(tree.data(), c.h_res + 64, sizeof(DPNode) * nodes_n);
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    st.gpu_ms = ms;
    if (profile) {
        for (int k = 2; k <= K; k++) {
            float b = 0.0f;
            HIPCHK(hipEventElapsedTime(&b, c.prof_ev[3 * k + 1], c.prof_ev[3 * k + 2]));
            st.eval_ms += b;
            st.eval_launches++;
        }
    }
    uint64_t made = 0;
    for (int k = 2; k <= K; k++) made += in.tree_counts[k];
    st.candidates = made;            /* sets generated by expansion (no 2^n scan) */
    st.sets = made;
    st.pairs = tree_pairs(in.tree_counts, (size_t)K);
    st.memo_bytes = total * (sizeof(float4) + 4) + list * 8;
    st.memo_kind = GQ_MEMO_RANK;
    return GQ_OK;
}

Tuning tuning_get() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    return g_tune;
}

int tuning_set(const Tuning &t) {
    if (t.lpe2_waves < 1 || t.lpe2_waves > (1u << 24) || t.gen_grid_max < 1 || t.gen_grid_max > 65536 ||
        t.tree_iters < 1 || t.tree_iters > 256 || t.expand_max_n > 32 || t.one_xcd_wgs > 256) {
        set_error("tuning out of range (two-lane waves 1..2^24, general grid 1..65536, tree task iterations "
                  "1..256, expansion up to 32 relations, one-XCD levels up to 256 workgroups)");
        return GQ_ERR_INVALID;
    }
    std::lock_guard<std::mutex> g(g_ctx.mu);
    std::lock_guard<std::mutex> g2(g_tune_mu);
    g_tune = t;
    return GQ_OK;
}

int resolve_memo(const DPInput &in, int requested, int max_memo_mb) {
    const int n = in.hdr.n_rels;
    const bool tree = in.hdr.is_tree && (int)in.tree_counts.size() == n + 1;
    if (!tree || requested == GQ_MEMO_DENSE) return GQ_MEMO_DENSE;
    if (requested == GQ_MEMO_HASH) return GQ_MEMO_HASH;
    if (requested == GQ_MEMO_RANK) return GQ_MEMO_RANK;
    const double dense = 36.0 * (double)(1ull << n);
    return dense > (double)max_memo_mb * 1048576.0 ? GQ_MEMO_RANK : GQ_MEMO_DENSE;
}

/* Rank tables of a tree graph whose vertex v > 0 has the lower-numbered
 * parent p(v) (BlobHdr::is_tree), so a connected set's lowest member is its
 * top.  Let N[v] be the number of connected sets with top v:
 * N[v] = prod over children c of (1 + N[c]).  The sets with top t are
 * numbered in mixed radix over t's children: child c (children by
 * increasing number) has weight M[c] = prod over its lower-numbered
 * siblings s of (1 + N[s]) and digit 0 (c absent) or 1 + the number of the
 * part below c (a set with top c).  These digits range over exactly
 * 1 + N[c] values, so the sets with top t get the numbers 0 .. N[t] - 1,
 * one each.  Unrolled, the number of T is
 *     sum over u in T - t of Q[t][u],  Q[t][u] = prod of M over the path
 * from t (exclusive) down to u,
 * and off[t] = sum of N[v] over v < t places the sets of each top after the
 * lower tops: every connected set gets a slot below their total.  The
 * blob gains off[n] and Q[n][n] (zero where t is not an ancestor of u).
 * A split at edge (p(c), c) of set S with top t gives the part below c,
 * sub, with slot off[c] + sig, sig = sum over u in sub - c of Q[c][u], and
 * since Q[t][u] = Q[t][c] Q[c][u] below c, the rest has slot
 * slot(S) - Q[t][c] (1 + sig).
 *
 * sig without a loop over sub: with w[u] = Q[0][u] (w[0] = 1), every u
 * below c has w[u] = w[c] Q[c][u], so W(sub) = sum over u in sub of w[u]
 * = w[c] (1 + sig).  W is linear in the bits of sub: four byte tables
 * B[j][v] = sum of w over the bits of v at byte j give it in four lookups.
 * W(sub) is below 2^31 (it is at most the slot of the set made of sub and
 * the path from 0 to c, and every slot is), so the u32 sums are exact, and
 * the exact quotient 1 + sig = W / w[c] is (W >> tz[c]) * inv[c] mod 2^32,
 * with w[c] = 2^tz[c] times an odd number whose inverse mod 2^32 is
 * inv[c].  The blob gains inv[n], tz[n] and B[ceil(n / 8)][256] after Q. */
static void attach_rank_tables(DPInput &in) {
    BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    std::vector<uint32_t> &b = in.blob;
    std::vector<int> parent(n, -1);
    std::vector<uint64_t> N(n, 1), M(n, 1);
    for (int v = 1; v < n; v++) parent[v] = __builtin_ctz(b[h.o_adj + v] & ((1u << v) - 1u));
    for (int v = n - 1; v >= 0; v--) {                 /* children are numbered above their parent */
        uint64_t prod = 1;
        for (int c = v + 1; c < n; c++)
            if (parent[c] == v) {
                M[c] = prod;
                prod *= 1 + N[c];
            }
        N[v] = prod;
    }
    const int o = (int)b.size();
    const int nbt = (n + 7) / 8;
    b.resize((size_t)o + n + (size_t)n * n + 2 * (size_t)n + 256 * (size_t)nbt, 0u);
    uint64_t acc = 0;
    for (int t = 0; t < n; t++) {
        b[o + t] = (uint32_t)acc;
        acc += N[t];
        uint32_t *q = b.data() + o + n + (size_t)t * n;
        for (int u = t + 1; u < n; u++) {
            if (parent[u] == t) q[u] = (uint32_t)M[u];
            else if (q[parent[u]]) q[u] = (uint32_t)((uint64_t)q[parent[u]] * M[u]);
        }
    }
    uint32_t *inv = b.data() + o + n + (size_t)n * n, *tz = inv + n, *bt = tz + n;
    for (int c = 0; c < n; c++) {
        const uint32_t w = c ? b[o + n + c] : 1u;     /* Q[0][c]: never 0, every vertex lies below 0 */
        tz[c] = (uint32_t)__builtin_ctz(w);
        const uint32_t odd = w >> tz[c];
        uint32_t x = odd;                               /* Newton: each step doubles the correct low bits */
        for (int i = 0; i < 5; i++) x *= 2u - odd * x;
        inv[c] = x;
        for (int v = 0; v < 256; v++)
            if ((v >> (c & 7)) & 1) bt[(c >> 3) * 256 + v] += w;
    }
    while (b.size() & 3) b.push_back(0u);
    h.o_rank = o;
    h.words = (int)b.size();
}

void prepare_memo(DPInput &in, int requested, int max_memo_mb) {
    in.memo = resolve_memo(in, requested, max_memo_mb);
    if (in.memo == GQ_MEMO_RANK) {
        uint64_t total = 0;
        for (size_t k = 1; k < in.tree_counts.size(); k++) total += in.tree_counts[k];
        /* slots and level-list offsets are 32-bit */
        if (total >= (1ull << 31)) in.memo = GQ_MEMO_HASH;
        else if (!in.hdr.o_rank) attach_rank_tables(in);
    }
}

/* ------------------------------------------------ memo snapshot (32-bit DPs)
 *
 * gq_plan_snapshot's copy of a finished DP's memo (dense, rank or open
 * addressing) into a table of 64-bit sets in the caller's numbering
 * (SnapTable, gq_wide.h): the connected sets are enumerated again level by
 * level, each from every set it extends (the first claim of its key keeps
 * it), and each takes its record and left side from the memo. */
struct Perm32 {
    int8_t p[32];
};

__device__ __forceinline__ uint64_t perm32(uint32_t x, const Perm32 &pm) {
    uint64_t r = 0;
    for (; x; x &= x - 1u) r |= 1ull << pm.p[__builtin_ctz(x)];
    return r;
}

/* cnt[k]: sets of size k copied; cnt[63]: error bits (1 table full or a
 * set twice, 2 list overflow, 4 a memo probe missed) */
__global__ void __launch_bounds__(WAVE) ks32_base(BlobHdr h, MemoRef mr, WTab to, Perm32 pm, uint32_t *lst,
                                                  uint32_t *cnt) {
    const int i = threadIdx.x;
    if (i >= h.n_rels) return;
    const uint32_t S = 1u << i;
    bool fresh;
    const uint32_t sl = winsert_fresh(to, perm32(S, pm), &fresh);
    if (sl == ~0u || !fresh) {
        atomicOr(&cnt[63], 1u);
        return;
    }
    const uint32_t ms = memo_slot(mr, S);
    to.s[sl].rec = mr.vals[ms];
    to.s[sl].left = 0ull;
    lst[i] = S;
    atomicAdd(&cnt[1], 1u);
}

__global__ void __launch_bounds__(BLOCK) ks32_level(BlobHdr h, const uint32_t *blob, MemoRef mr, WTab to,
                                                    Perm32 pm, const uint32_t *cur, uint32_t n_cur, uint32_t *nxt,
                                                    uint32_t cap, uint32_t *cnt, int k) {
    const uint32_t *adj = blob + h.o_adj;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n_cur; i += gridDim.x * BLOCK) {
        const uint32_t S = cur[i];
        uint32_t nb = 0;
        for (uint32_t x = S; x; x &= x - 1u) nb |= adj[__builtin_ctz(x)];
        for (uint32_t C = nb & ~S; C; C &= C - 1u) {
            const uint32_t T = S | (C & (0u - C));
            bool fresh;
            const uint32_t sl = winsert_fresh(to, perm32(T, pm), &fresh);
            if (sl == ~0u) {
                atomicOr(&cnt[63], 1u);
                continue;
            }
            if (!fresh) continue;
            const uint32_t pos = atomicAdd(&cnt[k + 1], 1u);
            if (pos >= cap) {
                atomicOr(&cnt[63], 2u);
                continue;
            }
            nxt[pos] = T;
            const uint32_t ms = memo_slot(mr, T);
            to.s[sl].rec = mr.vals[ms];
            to.s[sl].left = perm32(mr.left[ms], pm);
        }
    }
}

/* With the device context held, right after run_dp_locked planned `in`:
 * g_snap_req->out gets the memo's `total` connected sets. */
static int dp_snapshot(DeviceCtx &c, const DPInput &in, uint64_t total) {
    SnapReq &r = *g_snap_req;
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    hipStream_t s = c.stream;
    Perm32 pm;
    memcpy(pm.p, r.perm, 32);
    MemoRef mr = dense_memo(c);
    if (in.memo == GQ_MEMO_HASH) {
        uint64_t cap = 64;                         /* as run_dp_hash sized it */
        while (cap < 2 * total) cap <<= 1;
        mr.keys = c.hkeys;
        mr.mask = (uint32_t)(cap - 1);
    } else if (in.memo == GQ_MEMO_RANK) {
        mr.rank = c.b.blob + h.o_rank;
        mr.rn = n;
    }
    SnapTable &t = r.out;
    t.lg = snap_lg(total);
    t.n_rels = n;
    t.device = c.device;
    const size_t tbytes = sizeof(WSlot) << t.lg;
    uint32_t *lst = nullptr, *cnt = nullptr;
    if (hipMalloc(&t.tab, tbytes) != hipSuccess || hipMalloc((void **)&lst, 4 * total) != hipSuccess ||
        hipMalloc((void **)&cnt, 64 * 4) != hipSuccess) {
        (void)hipGetLastError();
        if (lst) (void)hipFree(lst);
        snapshot_free(t);
        set_error("hipMalloc for the memo snapshot failed");
        return GQ_ERR_NOMEM;
    }
    mr.err = cnt + 63;
    WTab to;
    to.s = (WSlot *)t.tab;
    to.lg = t.lg;
    to.mask = (uint32_t)((1ull << t.lg) - 1ull);
    uint32_t hc[64];
    int rc = GQ_OK;
    if (hipMemsetAsync(t.tab, 0, tbytes, s) != hipSuccess || hipMemsetAsync(cnt, 0, 64 * 4, s) != hipSuccess) {
        rc = GQ_ERR_DEVICE;
    } else {
        hipLaunchKernelGGL(ks32_base, dim3(1), dim3(WAVE), 0, s, h, mr, to, pm, lst, cnt);
        uint64_t off = 0, done = (uint64_t)n;
        for (int k = 1; k < n && rc == GQ_OK; k++) {
            if (hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = GQ_ERR_DEVICE;
                break;
            }
            if (hc[63]) break;
            const uint32_t nk = hc[k];
            const uint64_t nxt = off + nk;
            if (nk == 0 || nxt > total) {
                hc[63] |= 2u;
                break;
            }
            hipLaunchKernelGGL(ks32_level, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(
                                             ((uint64_t)nk + BLOCK - 1) / BLOCK, 4096))),
                               dim3(BLOCK), 0, s, h, (const uint32_t *)c.b.blob, mr, to, pm, lst + off, nk,
                               lst + nxt, (uint32_t)(total - nxt), cnt, k);
            off = nxt;
        }
        if (rc == GQ_OK && (hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s) != hipSuccess ||
                            hipStreamSynchronize(s) != hipSuccess))
            rc = GQ_ERR_DEVICE;
        if (rc == GQ_OK) {
            done = 0;
            for (int k = 1; k <= n; k++) done += hc[k];
            if (hc[63] || done != total) {
                set_error("memo snapshot: %llu of %llu connected sets copied (error bits %u)",
                          (unsigned long long)done, (unsigned long long)total, hc[63]);
                rc = GQ_ERR_DEVICE;
            }
        } else {
            set_error("memo snapshot: a device call failed");
        }
    }
    (void)hipFree(lst);
    (void)hipFree(cnt);
    if (rc) {
        snapshot_free(t);
        return rc;
    }
    t.n_sets = total;
    r.done = true;
    return GQ_OK;
}

static int run_dp_locked(const DPInput &in, DPAlgo algo, int n_iters, int memo_kind, bool profile,
                         std::vector<DPNode> &tree, DPStats &st, bool noalloc);

int run_dp(const DPInput &in, DPAlgo algo, int n_iters, int memo_kind, bool profile, std::vector<DPNode> &tree,
           DPStats &st, bool noalloc) {
    std::lock_guard<std::mutex> g(g_ctx.mu);
    int rc = run_dp_locked(in, algo, n_iters, memo_kind, profile, tree, st, noalloc);
    if (rc == GQ_OK && g_snap_req && g_snap_req->n == in.hdr.n_rels) {
        if (noalloc || algo == DP_DPSIZE || (n_iters >= 2 && n_iters < in.hdr.n_rels)) {
            set_error("a memo snapshot takes one DPsub/MPDP DP without a level cap");
            return GQ_ERR_UNSUPPORTED;
        }
        rc = dp_snapshot(g_ctx, in, st.sets + (uint64_t)in.hdr.n_rels);
    }
    return rc;
}

static int run_dp_locked(const DPInput &in, DPAlgo algo, int n_iters, int memo_kind, bool profile,
                         std::vector<DPNode> &tree, DPStats &st, bool noalloc) {
    DeviceCtx &c = g_ctx;
    int rc = ctx_init(c, c.device < 0 ? 0 : c.device);
    if (rc) return rc;
    const bool counters_were_clean = c.counters_clean && !noalloc;
    if (!noalloc) c.counters_clean = false;
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    if (n < 1 || n > 32) {
        set_error("GPU DP supports 1..32 relations, got %d", n);
        return GQ_ERR_UNSUPPORTED;
    }
    /* n_iters < n: IDP1's level cap; the result is the cheapest set of
     * n_iters relations */
    const int K = (n_iters >= 2 && n_iters < n) ? n_iters : n;
    const int mode = dp_mode(in, algo);
    if (noalloc) {
        /* everything below runs in `scratch`: one dense DP, no IDP1 cap */
        if (n > GQ_NOALLOC_MAX_RELS || algo == DP_DPSIZE || K != n || in.memo != GQ_MEMO_DENSE) {
            set_error("the allocation-free path plans dense-memo DPsub/MPDP DPs of at most %d relations",
                      GQ_NOALLOC_MAX_RELS);
            return GQ_ERR_UNSUPPORTED;
        }
        profile = false;
    }
    if (in.memo == GQ_MEMO_HASH && algo != DP_DPSIZE) return run_dp_hash(c, in, algo, K, profile, tree, st);
    if (in.memo == GQ_MEMO_RANK && algo != DP_DPSIZE) return run_dp_rank(c, in, algo, K, profile, tree, st);
    const size_t blob_bytes = (size_t)h.words * 4;
    const size_t lds_eval = ((blob_bytes + 15) & ~(size_t)15);
    const uint32_t conn_words_all = (uint32_t)std::max<uint64_t>(1, (1ull << n) / 32);
    const int conn_lds = conn_words_all <= 4096 ? (int)conn_words_all : 0;
    const size_t lds_bicc = eval_lds_bicc(h.words);
    const size_t lds_dpsub = eval_lds_dpsub(h.words, conn_lds);
    if (lds_bicc > 64 * 1024 || lds_dpsub > 64 * 1024) {
        set_error("query description too large for LDS (%zu bytes)", blob_bytes);
        return GQ_ERR_UNSUPPORTED;
    }
    (void)memo_kind;
    const uint64_t nsets_all = 1ull << n;
    uint64_t maxC = 1;
    for (int k = 1; k <= n; k++) maxC = std::max(maxC, binom64(n, k));
    DevBufs nb{};
    BigEnt *big_ents = nullptr;
    uint32_t *big_prefix = nullptr;
    if (noalloc) {
        if ((rc = scratch_bufs(c.device, nb, &big_ents, &big_prefix))) return rc;
    } else {
        if ((rc = ensure(&c.b.vals, c.cap_memo, nsets_all))) return rc;
        if ((rc = ensure(&c.b.left, c.cap_left, nsets_all))) return rc;
        if ((rc = ensure(&c.b.conn, c.cap_conn, std::max<uint64_t>(1, nsets_all / 32)))) return rc;
        if ((rc = ensure(&c.b.lvl_sets, c.cap_lvl, nsets_all))) return rc;
        if ((rc = ensure(&c.b.lvl_best, c.cap_best, nsets_all))) return rc;
        if ((rc = ensure(&c.b.lvl_first, c.cap_first, nsets_all))) return rc;
    }
    /* device blob area: the query blob, then (at a 4-word boundary) the 33
     * level-list offsets; both go up in one copy */
    const size_t W4h = ((size_t)h.words + 3) & ~(size_t)3;
    /* staged area: blob | level offsets (33, padded to 36) | DPsize level-1
     * order (32) | the big-set job descriptor */
    const size_t o_l1 = W4h + 36, o_sj = o_l1 + 32;
    const size_t stage_words = o_sj + ((sizeof(BatchJob) + 15) & ~(size_t)15) / 4;
    if (noalloc) {
        if (stage_words > na::STAGE_WORDS || c.cap_h_stage < stage_words || maxC > na::MAXC) {
            set_error("query description too large for the allocation-free path");
            return GQ_ERR_UNSUPPORTED;
        }
    } else {
        if ((rc = ensure(&c.b.blob, c.cap_blob, stage_words))) return rc;
        if ((rc = ensure_stage(c, stage_words))) return rc;
    }
    /* the DP's buffers from here on: the context's, or `scratch` */
    DevBufs &b = noalloc ? nb : c.b;
    uint32_t *const lvl_off = b.blob + W4h;

    /* optional per-phase events: [3k] before filter, [3k+1] before eval, [3k+2] after eval */
    if (profile && (int)c.prof_ev.size() < 3 * 33) {
        while ((int)c.prof_ev.size() < 3 * 33) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            c.prof_ev.push_back(e);
        }
    }
    hipStream_t s = c.stream;
    /* level lists: level k at offset sum_{2<=j<k} C(n,j) */
    uint32_t off_h[33];
    memset(off_h, 0, sizeof(off_h));
    {
        uint64_t acc = 0;
        for (int k = 2; k <= n; k++) {
            off_h[k] = (uint32_t)acc;
            acc += binom64(n, k);
        }
    }
    memset(c.h_stage, 0, stage_words * 4);
    memcpy(c.h_stage, in.blob.data(), blob_bytes);
    memcpy(c.h_stage + W4h, off_h, sizeof(off_h));
    if (algo == DP_DPSIZE) memcpy(c.h_stage + o_l1, in.level1.data(), 4 * std::min<size_t>(in.level1.size(), 32));
    const uint32_t conn_words = (uint32_t)std::max<uint64_t>(1, nsets_all / 32);
    /* MPDP on a graph with cycles of 15 or more relations: sets with a big
     * block are deferred to the chunk-parallel big-set kernels, which see
     * this query as a batch of one. */
    const bool use_big = algo == DP_BICC && !h.is_tree && n >= 15;
    BatchJob *const d_sj = (BatchJob *)(b.blob + o_sj);
    if (use_big) {
        BatchJob &sj = *(BatchJob *)(c.h_stage + o_sj);
        sj.h = h;
        sj.blob = b.blob;
        sj.vals = b.vals;
        sj.left = b.left;
        sj.conn = b.conn;
        sj.lvl_sets = b.lvl_sets;
        sj.counters = b.counters;
        sj.lvl_best = b.lvl_best;
        sj.lvl_first = b.lvl_first;
        sj.nodes = b.nodes;
        memcpy(sj.lvl_off, off_h, sizeof(off_h));
        sj.conn_words = conn_words;
        sj.bicc = 1;
    }
    /* Trees with exact level sizes and at most expand_max_n relations skip
     * the 2^n filter: level 2 comes from k_tree_start and each level's
     * launch builds the next level's list by expansion (no connectivity
     * bitmap is needed).  The limit is a tuning knob: the autotuner times
     * both producers on the caller's sample (on the 20-relation snowflake
     * expansion is faster; above the limit the filter's single pass is
     * kept). */
    const bool tree_exp = h.is_tree && (int)in.tree_counts.size() == n + 1 && algo != DP_DPSIZE && n >= 2 &&
                          n <= (int)g_tune.expand_max_n;
    /* fused: level 2 starts the DP (TREE_START) and the last level walks the
     * plan into pinned host memory (TREE_WALK), so no start kernel, walk
     * kernel or read-back copy runs.  Unprofiled, the host then waits for the
     * walk's status word (wait_status) and takes the device time from the
     * device clock the first and last launches read, not from events. */
    const bool fused = tree_exp && K == n;
    const bool poll = fused && !profile;
    if (poll) c.h_res[48] = RES_PENDING;
    st.t_first_launch = host_now_ms();
    if ((rc = stage_up(c, s, b.blob, stage_words))) return rc;
    if (!poll) HIPCHK(hipEventRecord(c.ev0, s));
    if (fused) {
        if (!counters_were_clean) hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, s, nullptr, 0u, b.counters);
    } else if (tree_exp) {
        hipLaunchKernelGGL(k_tree_start, dim3(1), dim3(WAVE), 0, s, h, b.blob, b.vals, b.left,
                           b.lvl_sets + off_h[2], b.counters, nullptr);
    } else {
        uint32_t blocks = (uint32_t)std::min<uint64_t>((conn_words + BLOCK - 1) / BLOCK, 4096);
        hipLaunchKernelGGL(k_clear, dim3(blocks), dim3(BLOCK), 0, s, b.conn, conn_words, b.counters);
        hipLaunchKernelGGL(k_base, dim3(1), dim3(64), 0, s, h, b.blob, b.vals, b.left, b.conn);
    }
    if (algo == DP_DPSIZE) {
        if (n > 28) {
            set_error("GPU DPsize supports at most 28 relations (got %d)", n);
            return GQ_ERR_UNSUPPORTED;
        }
        if ((int)in.level1.size() != n) {
            set_error("DPsize level-1 order missing");
            return GQ_ERR_INVALID;
        }
        DsBufs &d = c.ds;
        if ((rc = ensure(&d.best, c.cap_ds_best, nsets_all))) return rc;
        if ((rc = ensure(&d.first, c.cap_ds_first, nsets_all))) return rc;
        if ((rc = ensure(&d.lset, c.cap_ds_lset, nsets_all))) return rc;
        if ((rc = ensure(&d.lnbr, c.cap_ds_lnbr, nsets_all))) return rc;
        if ((rc = ensure(&d.raw, c.cap_ds_raw, maxC))) return rc;
        if ((rc = ensure(&d.tmp, c.cap_ds_tmp, maxC))) return rc;
        if ((rc = ensure(&d.rowcnt, c.cap_ds_rc, nsets_all + 1))) return rc;
        if ((rc = ensure(&d.rowfill, c.cap_ds_rf, nsets_all + 1))) return rc;
        if ((rc = ensure(&d.rowoff, c.cap_ds_ro, nsets_all + 1))) return rc;
        if ((rc = ensure(&d.sz, c.cap_ds_sz, 256))) return rc;

        const size_t lds_pairs = lds_eval + (DS_TA * 2 + DS_TB + DS_TA * DS_TB + 4) * 4;
        {
            uint32_t blocks = (uint32_t)std::min<uint64_t>((nsets_all + BLOCK - 1) / BLOCK, 4096);
            hipLaunchKernelGGL(k_ds_init, dim3(blocks), dim3(BLOCK), 0, s, h, b.blob, (const int *)(b.blob + o_l1),
                               d, nsets_all, b.vals,
                               b.left, b.counters);
        }
        uint64_t cand_ub = 0;
        LevelGate gate{c, s, paced_dp(in, algo)};
        for (int k = 2; k <= K; k++) {
            if ((rc = gate.before(k))) return rc;
            if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k], s));
            hipLaunchKernelGGL(k_ds_prep, dim3(1), dim3(64), 0, s, k, n, d);
            uint32_t rblocks = (uint32_t)std::min<uint64_t>((nsets_all + BLOCK) / BLOCK, 4096);
            hipLaunchKernelGGL(k_ds_zero_rows, dim3(rblocks), dim3(BLOCK), 0, s, d);
            /* grid: upper bound on the tiles of this level from binomials */
            uint64_t tiles_ub = 0;
            for (int t = 0; t <= k - 2; t++) {
                uint64_t nl = binom64(n, k - 1 - t), nr = binom64(n, t + 1);
                tiles_ub += ((nl + DS_TA - 1) / DS_TA) * ((nr + DS_TB - 1) / DS_TB);
                cand_ub += nl * nr;
            }
            uint32_t pblocks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(tiles_ub, 1), 8192);
            if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 1], s));
            hipLaunchKernelGGL(k_ds_pairs, dim3(pblocks), dim3(BLOCK), lds_pairs, s, h, b.blob, k, d, b.vals,
                               b.counters);
            uint64_t Ck = binom64(n, k);
            uint32_t sblocks = (uint32_t)std::min<uint64_t>((Ck + BLOCK - 1) / BLOCK, 4096);
            hipLaunchKernelGGL(k_ds_rowcount, dim3(sblocks), dim3(BLOCK), 0, s, k, d, b.counters);
            hipLaunchKernelGGL(k_ds_scan, dim3(1), dim3(1024), 0, s, d);
            hipLaunchKernelGGL(k_ds_scatter, dim3(sblocks), dim3(BLOCK), 0, s, k, d, b.counters);
            hipLaunchKernelGGL(k_ds_rowsort, dim3(rblocks), dim3(BLOCK), 0, s, d);
            hipLaunchKernelGGL(k_ds_finalize, dim3(sblocks), dim3(BLOCK), lds_eval, s, h, b.blob, k, d, b.vals,
                               b.left, b.counters);
            if (profile) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 2], s));
            if ((rc = gate.after(k))) return rc;
        }
        (void)cand_ub;
        uint32_t root = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
        if (K < n) {                               /* IDP1: the cheapest set of K relations */
            /* creation order: cpu_dpsize's level lists (gpuqo_cpu_dpsize.cu:55-92), which d.lset reproduces */
            auto created = [&](std::vector<uint32_t> &o) -> int {
                std::vector<uint32_t> sz(256);
                HIPCHK(hipMemcpy(sz.data(), d.sz, 256 * 4, hipMemcpyDeviceToHost));
                const uint64_t total = (uint64_t)sz[40 + K] + sz[K];
                std::vector<uint32_t> all(total);
                if (total) HIPCHK(hipMemcpy(all.data(), d.lset, total * 4, hipMemcpyDeviceToHost));
                o.clear();
                for (int k = 2; k <= K; k++) o.insert(o.end(), all.begin() + sz[40 + k], all.begin() + sz[40 + k] + sz[k]);
                return (int)GQ_OK;
            };
            rc = idp1_root(c, s, dense_memo_of(b), d.lset, d.sz + 40 + K, d.sz + K, n, K, created, &root);
            if (rc) return rc;
        }
        const int nodes_n = 2 * K - 1;
        hipLaunchKernelGGL(k_tree, dim3(1), dim3(64), 0, s, dense_memo_of(b), root, b.nodes, nodes_n,
                           b.counters + 48);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c.ev1, s));
        tree.resize(nodes_n);
        const uint32_t *cnt = c.h_res;
        st.t_last_launch = host_now_ms();
        if ((rc = results_down(c, s, nodes_n, b.counters))) return rc;
        HIPCHK(hipStreamSynchronize(s));
        st.t_synced = host_now_ms();
        memcpy(tree.data(), c.h_res + 64, sizeof(DPNode) * nodes_n);
        if (cnt[48] != (uint32_t)nodes_n) {
            set_error("plan reconstruction failed (memo inconsistent, status %u)", cnt[48]);
            return GQ_ERR_DEVICE;
        }
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        st.gpu_ms = ms;
        if (profile) {
            for (int k = 2; k <= K; k++) {
                float a = 0.0f, b = 0.0f;
                HIPCHK(hipEventElapsedTime(&a, c.prof_ev[3 * k], c.prof_ev[3 * k + 1]));
                HIPCHK(hipEventElapsedTime(&b, c.prof_ev[3 * k + 1], c.prof_ev[3 * k + 2]));
                st.filter_ms += a;
                st.eval_ms += b;
                st.filter_launches++;
                st.eval_launches++;
            }
        }
        st.candidates = (uint64_t)cnt[44] | ((uint64_t)cnt[45] << 32);
        st.sets = 0;
        for (int k = 2; k <= K; k++) st.sets += cnt[k];
        st.pairs = (uint64_t)cnt[40] | ((uint64_t)cnt[41] << 32);
        st.memo_bytes = nsets_all * (sizeof(float4) + 4 + 16 + 8 + 12);
        st.memo_kind = 1;
        return GQ_OK;
    }
    if (use_big) {
        if (!noalloc) {
            if ((rc = ensure(&c.big_ents, c.cap_big, maxC))) return rc;
            if ((rc = ensure(&c.big_prefix, c.cap_big_prefix, maxC + 1))) return rc;
            big_ents = c.big_ents;
            big_prefix = c.big_prefix;
        }
    }
    const uint64_t cand_total = nsets_all - (uint64_t)n - 1ull;
    if (profile) HIPCHK(hipEventRecord(c.prof_ev[0], s));
    if (!tree_exp) {
        uint64_t waves = (nsets_all + WAVE - 1) / WAVE;
        uint32_t fblocks = (uint32_t)std::min<uint64_t>((waves + 3) / 4, 8192);
        hipLaunchKernelGGL(k_filter_all, dim3(fblocks), dim3(BLOCK), 0, s, h, b.blob, lvl_off, b.lvl_sets,
                           b.lvl_best, b.lvl_first, b.conn, conn_words, b.counters);
    }
    if (profile) HIPCHK(hipEventRecord(c.prof_ev[1], s));
    LevelGate gate{c, s, paced_dp(in, algo)};
    for (int k = 2; k <= K; k++) {
        if ((rc = gate.before(k))) return rc;
        uint64_t Ck = binom64(n, k);
        uint32_t *sets_k = b.lvl_sets + off_h[k];
        unsigned long long *best_k = b.lvl_best + off_h[k];
        uint32_t *first_k = b.lvl_first + off_h[k];
        uint32_t lblocks = (uint32_t)std::min<uint64_t>((Ck + BLOCK - 1) / BLOCK, 8192);
        /* trees: one launch per level, timed as one span from the first
         * level to the last (events between launches would add their own
         * gaps to every level) */
        if (profile && (!h.is_tree || k == 2)) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 1], s));
        const uint64_t T = (1ull << k) - 2;
        if (h.is_tree) {
            /* exact level size of a tree (subtree polynomial), else C(n,k) */
            const uint64_t lvl = (int)in.tree_counts.size() == n + 1 ? in.tree_counts[k] : Ck;
            const uint32_t tf = fused ? ((k == 2 ? TREE_START : 0u) | (k == K ? TREE_WALK : 0u)) : 0u;
            launch_eval_tree(h, lvl, lds_eval, s, b.blob, k, algo == DP_BICC ? 1 : 0, sets_k, nullptr,
                             dense_memo_of(b), b.counters,
                             tree_exp && k < K ? b.lvl_sets + off_h[k + 1] : nullptr, nullptr, tf, c.h_res);
        } else if (mode != 0 || T <= 4096) {
            uint64_t items_est = (mode != 0 || h.is_tree) ? 2ull * (k - 1) : T;
            uint32_t B = mode != 0 ? task_sets_of(false, true, k, g_tune.tree_iters)
                                   : (uint32_t)std::min<uint64_t>(BMAX, std::max<uint64_t>(1, (768 + items_est - 1) / items_est));
            uint32_t wblocks = (uint32_t)std::min<uint64_t>((Ck + B - 1) / B, 4096);
            BigList bl;
            bl.ents = nullptr;
            bl.count = nullptr;
            bl.cap = 0;
            bl.job = 0;
            bl.gidx0 = 0;
            if (algo == DP_BICC) {
                const bool big = use_big && k >= 15;
                if (big) {
                    bl.ents = big_ents;
                    bl.count = b.counters + 50;
                    bl.cap = (uint32_t)maxC;
                    bl.gidx0 = off_h[k];
                }
                hipLaunchKernelGGL(k_eval_wg<true>, dim3(wblocks), dim3(BLOCK), lds_bicc, s, h, b.blob, k, B, 0,
                                   sets_k, b.vals, b.left, b.conn, b.counters, bl);
                if (big) {
                    hipLaunchKernelGGL(k_big_scan, dim3(1), dim3(1024), 0, s, big_ents, b.counters + 50,
                                       (uint32_t)maxC, big_prefix);
                    hipLaunchKernelGGL(k_big_eval, dim3(1024), dim3(BLOCK), big_lds(h.words), s, d_sj, big_ents,
                                       b.counters + 50, big_prefix);
                }
            } else if (mode == 2) {          /* DPsub splits through the blocks (dpsub_by_blocks) */
                hipLaunchKernelGGL((k_eval_wg<true, true>), dim3(wblocks), dim3(BLOCK), lds_bicc, s, h, b.blob, k,
                                   B, 0, sets_k, b.vals, b.left, b.conn, b.counters, bl);
            } else {
                hipLaunchKernelGGL(k_eval_wg<false>, dim3(wblocks), dim3(BLOCK), lds_dpsub, s, h, b.blob, k, B,
                                   conn_lds, sets_k, b.vals, b.left, b.conn, b.counters, bl);
            }
        } else {
            uint32_t ipl = 16;
            uint64_t per_wave = (uint64_t)ipl * WAVE;
            uint32_t cps = (uint32_t)((T + per_wave - 1) / per_wave);
            uint64_t waves = Ck * cps;
            uint32_t wblocks = (uint32_t)std::min<uint64_t>((waves + 3) / 4, 16384);
            hipLaunchKernelGGL(k_eval_dpsub_chunk, dim3(wblocks), dim3(BLOCK), lds_eval, s, h, b.blob, k, cps, ipl,
                               sets_k, best_k, first_k, b.vals, b.conn, b.counters);
            hipLaunchKernelGGL(k_finalize_dpsub, dim3(lblocks), dim3(BLOCK), lds_eval, s, h, b.blob, k, sets_k,
                               best_k, first_k, b.vals, b.left, b.counters);
        }
        if (profile && (!h.is_tree || k == K)) HIPCHK(hipEventRecord(c.prof_ev[3 * k + 2], s));
        if ((rc = gate.after(k))) return rc;
    }
    uint32_t root = (n == 32) ? 0xffffffffu : ((1u << n) - 1u);
    if (K < n) {                                   /* IDP1: the cheapest set of K relations */
        rc = idp1_root(c, s, dense_memo_of(b), b.lvl_sets + off_h[K], nullptr, b.counters + K, n, K,
                       [&](std::vector<uint32_t> &o) -> int { o = dpsub_creation_order(in, K); return GQ_OK; }, &root);
        if (rc) return rc;
    }
    const int nodes_n = 2 * K - 1;
    if (!fused) hipLaunchKernelGGL(k_tree, dim3(1), dim3(64), 0, s, dense_memo_of(b), root, b.nodes, nodes_n,
                                   b.counters + 48);
    HIPCHK(hipGetLastError());
    if (!poll) HIPCHK(hipEventRecord(c.ev1, s));
    tree.resize(nodes_n);
    const uint32_t *cnt = c.h_res;
    st.t_last_launch = host_now_ms();
    if (!fused && (rc = results_down(c, s, nodes_n, b.counters))) return rc;
    if (poll) {
        if ((rc = wait_status(c, s))) return rc;
    } else {
        HIPCHK(hipStreamSynchronize(s));
    }
    st.t_synced = host_now_ms();
    if (fused && !noalloc) c.counters_clean = true;   /* the walk zeroed them */
    memcpy(tree.data(), c.h_res + 64, sizeof(DPNode) * nodes_n);
    if (cnt[49] != 0) {
        set_error("a connected set has more than 2^31 splits (unsupported on the GPU path)");
        return GQ_ERR_UNSUPPORTED;
    }
    if (cnt[48] != (uint32_t)nodes_n) {
        set_error("plan reconstruction failed (memo inconsistent, status %u)", cnt[48]);
        return GQ_ERR_DEVICE;
    }
    float ms = 0.0f;
    if (poll) {
        const uint64_t t0 = (uint64_t)cnt[56] | ((uint64_t)cnt[57] << 32);
        const uint64_t t1 = (uint64_t)cnt[58] | ((uint64_t)cnt[59] << 32);
        ms = c.wall_khz > 0 && t1 >= t0 ? (float)((double)(t1 - t0) / (double)c.wall_khz) : 0.0f;
    } else {
        HIPCHK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    }
    st.gpu_ms = ms;
    if (profile) {
        float a = 0.0f;
        HIPCHK(hipEventElapsedTime(&a, c.prof_ev[0], c.prof_ev[1]));
        st.filter_ms = a;
        st.filter_launches = 1;
        if (h.is_tree && K >= 2) {
            float b = 0.0f;
            HIPCHK(hipEventElapsedTime(&b, c.prof_ev[3 * 2 + 1], c.prof_ev[3 * K + 2]));
            st.eval_ms = b;
            st.eval_launches = K - 1;
        } else {
            for (int k = 2; k <= K; k++) {
                float b = 0.0f;
                HIPCHK(hipEventElapsedTime(&b, c.prof_ev[3 * k + 1], c.prof_ev[3 * k + 2]));
                st.eval_ms += b;
                st.eval_launches++;
            }
        }
    }
    st.candidates = cand_total;
    if (tree_exp) {                        /* no 2^n scan: the level lists were built by expansion */
        st.candidates = 0;
        for (int k = 2; k <= K; k++) st.candidates += in.tree_counts[k];
    }
    st.sets = 0;
    for (int k = 2; k <= K; k++) st.sets += cnt[k];
    st.pairs = (uint64_t)cnt[40] | ((uint64_t)cnt[41] << 32);
    if (h.is_tree && (int)in.tree_counts.size() == n + 1) st.pairs = tree_pairs(in.tree_counts, (size_t)K);
    st.memo_bytes = nsets_all * (sizeof(float4) + 4);
    st.memo_kind = 1;
    return GQ_OK;
}


/* ================================================================ step engine
 *
 * Continuous batching over DP levels.  An admitted DP gets one slice of the
 * engine's device arena (256-byte aligned):
 *     [result record: 64 counters + 64 plan nodes][query blob]
 *     [vals][left][conn][lvl_sets][lvl_best][lvl_first]
 * On its first step the DP is initialised and filtered; each step then runs
 * one level of every active DP in one k_batch_eval launch (each at its own
 * level J.cur_k) plus the big-block kernels when a DP needs them; a DP past
 * its last level is walked and read back, and its slice is freed.  The host
 * knows every DP's level, so no device round trip decides the schedule; one
 * stream synchronisation per step returns the finished plans. */

DPExecutor *&thread_executor() {
    static thread_local DPExecutor *ex = nullptr;
    return ex;
}

namespace {
inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
const size_t ENGINE_BLOB_MAX = 64 * 1024;     /* blob bytes allowed per DP (the LDS bound) */

size_t slice_layout(int n, int words, size_t *o_blob, size_t *o_vals, size_t *o_left, size_t *o_conn,
                    size_t *o_sets, size_t *o_best, size_t *o_first) {
    const uint64_t ns = 1ull << n;
    size_t o = al256(RES_WORDS * 4);
    *o_blob = o;
    o += al256((size_t)words * 4);
    *o_vals = o;
    o += al256(ns * 16);
    *o_left = o;
    o += al256(ns * 4);
    *o_conn = o;
    o += al256(std::max<uint64_t>(1, ns / 32) * 4);
    *o_sets = o;
    o += al256(ns * 4);
    *o_best = o;
    o += al256(ns * 8);
    *o_first = o;
    o += al256(ns * 4);
    return o;
}
}  // namespace

struct EJob {
    uint64_t key;
    int n, cur_k;
    DPAlgo algo;
    size_t off, bytes;
    BatchJob J;
    int words;
    bool hash;
    bool expand;                         /* level lists by k_expand (hash, rank, dense trees): counts exact */
    uint64_t cap;                        /* open-addressing capacity */
    std::vector<uint64_t> counts;        /* connected sets per size (tree graphs) */
    std::vector<uint32_t> blob;          /* query blob, copied to the device by the DP's first step */
    int memo_kind = GQ_MEMO_DENSE;
};

namespace {
/* open-addressing slice: [result][blob][keys][vals][left][lvl_sets][lvl_slot];
 * rank slice (rank = true): [result][blob][vals][left][lvl_sets][lvl_slot],
 * one memo slot per connected set */
size_t hash_slice_layout(uint64_t total_sets, uint64_t list, int words, bool rank, uint64_t *cap, size_t *o_blob,
                         size_t *o_keys, size_t *o_vals, size_t *o_left, size_t *o_sets, size_t *o_slot) {
    uint64_t c = 64;
    while (c < 2 * total_sets) c <<= 1;
    if (rank) c = std::max<uint64_t>(total_sets, 1);
    *cap = c;
    size_t o = al256(RES_WORDS * 4);
    *o_blob = o;
    o += al256((size_t)words * 4);
    *o_keys = o;
    if (!rank) o += al256(c * 4);
    *o_vals = o;
    o += al256(c * 16);
    *o_left = o;
    o += al256(c * 4);
    *o_sets = o;
    o += al256(std::max<uint64_t>(list, 1) * 4);
    *o_slot = o;
    o += al256(std::max<uint64_t>(list, 1) * 4);
    return o;
}
}  // namespace

uint64_t engine_hash_slice_bytes(uint64_t total_sets) {
    uint64_t cap;
    size_t a, b, c, d, e, f;
    return hash_slice_layout(total_sets, total_sets, (int)(ENGINE_BLOB_MAX / 4), false, &cap, &a, &b, &c, &d, &e, &f);
}

uint64_t engine_rank_slice_bytes(uint64_t total_sets) {
    uint64_t cap;
    size_t a, b, c, d, e, f;
    return hash_slice_layout(total_sets, total_sets, (int)(ENGINE_BLOB_MAX / 4), true, &cap, &a, &b, &c, &d, &e, &f);
}

/* One step's host side.  Steps alternate between two slots: the engine
 * enqueues step i (slot i % 2) before it waits for step i - 1, so the host
 * prepares and launches one step while the device runs the previous one.
 * A slot's pinned packet and result buffer are rewritten only after its
 * previous step has completed. */
struct StepSlot {
    /* pinned, coherent host memory that the step's kernels read and write
     * directly: no copy engine between the host and the step */
    uint8_t *h_pkt = nullptr;                    /* the step's packet, staged by its first kernel */
    uint32_t *h_res = nullptr;                   /* RES_WORDS per finishing DP, written by the walk */
    size_t cap_h_pkt = 0, cap_h_res = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<EJob *> fin;                     /* DPs walked in this step, read back at its completion */
    bool busy = false;
};

struct StepEngine {
    int device = 0;
    hipStream_t s = nullptr, s2 = nullptr;       /* tree jobs / general jobs (default priority) */
    hipEvent_t fork = nullptr, join = nullptr;
    char *arena = nullptr;
    size_t arena_bytes = 0;
    std::map<size_t, size_t> free_;              /* offset -> size, coalesced */
    std::vector<EJob *> active, incoming;
    uint8_t *d_pkt = nullptr;                    /* device copy of the current step's packet */
    uint32_t *d_tp = nullptr, *big_prefix = nullptr, *big_hdr = nullptr;
    BigEnt *big = nullptr;
    size_t cap_pkt = 0, cap_tp = 0, cap_big = 0, cap_bigp = 0, cap_hdr = 0;
    StepSlot slot[2];
    int cur = 0;                                 /* slot of the next step */

    bool alloc(size_t bytes, size_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

template <class T>
static int ensure_pinned(T **p, size_t &cap, size_t need, unsigned flags = hipHostMallocDefault) {
    if (*p && cap >= need) return GQ_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    const size_t n = std::max<size_t>(std::max<size_t>(need, 16), 2 * cap);
    cap = 0;
    if (hipHostMalloc((void **)p, n * sizeof(T), flags) != hipSuccess) {
        *p = nullptr;
        set_error("pinned allocation of %zu bytes failed", n * sizeof(T));
        return GQ_ERR_NOMEM;
    }
    cap = n;
    return GQ_OK;
}

uint64_t engine_slice_bytes(int n_rels) {
    size_t a, b, c, d, e, f, g;
    return slice_layout(n_rels, (int)(ENGINE_BLOB_MAX / 4), &a, &b, &c, &d, &e, &f, &g);
}

int engine_create(int device, uint64_t arena_bytes, StepEngine **out) {
    *out = nullptr;
    StepEngine *e = new StepEngine();
    e->device = device;
    auto fail = [&](int rc) {
        engine_destroy(e);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) {
        set_error("hipSetDevice(%d) failed", device);
        return fail(GQ_ERR_DEVICE);
    }
    {
        /* the binomial table and device context of the library */
        int rc = device_init(device);
        if (rc) return fail(rc);
    }
    bool ok = hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&e->s2, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&e->join, hipEventDisableTiming) == hipSuccess;
    for (StepSlot &sl : e->slot)
        ok = ok && hipEventCreate(&sl.e0) == hipSuccess && hipEventCreate(&sl.e1) == hipSuccess;
    if (!ok) {
        set_error("engine stream/event creation failed");
        return fail(GQ_ERR_DEVICE);
    }
    e->arena_bytes = al256(arena_bytes);
    if (hipMalloc((void **)&e->arena, e->arena_bytes) != hipSuccess) {
        e->arena = nullptr;
        set_error("engine arena of %llu bytes could not be allocated", (unsigned long long)e->arena_bytes);
        return fail(GQ_ERR_NOMEM);
    }
    e->free_[0] = e->arena_bytes;
    int rc = ensure(&e->big_hdr, e->cap_hdr, 4);
    if (rc) return fail(rc);
    /* staging sized for a few hundred DPs up front: growing pinned memory
     * in the middle of a burst of requests costs milliseconds */
    const unsigned host_flags = hipHostMallocMapped | hipHostMallocCoherent;
    for (StepSlot &sl : e->slot) {
        if ((rc = ensure_pinned(&sl.h_pkt, sl.cap_h_pkt, (size_t)1 << 20, host_flags))) return fail(rc);
        if ((rc = ensure_pinned(&sl.h_res, sl.cap_h_res, (size_t)128 * RES_WORDS, host_flags))) return fail(rc);
    }
    if ((rc = ensure(&e->d_pkt, e->cap_pkt, (size_t)1 << 20))) return fail(rc);
    if ((rc = ensure(&e->d_tp, e->cap_tp, (size_t)2 * 1025))) return fail(rc);
    if (hipMemsetAsync(e->big_hdr, 0, 16, e->s) != hipSuccess || hipStreamSynchronize(e->s) != hipSuccess) {
        set_error("engine initialisation failed");
        return fail(GQ_ERR_DEVICE);
    }
    *out = e;
    return GQ_OK;
}

void engine_destroy(StepEngine *e) {
    if (!e) return;
    if (e->s) (void)hipStreamSynchronize(e->s);
    if (e->s2) (void)hipStreamSynchronize(e->s2);
    for (EJob *j : e->active) delete j;
    for (EJob *j : e->incoming) delete j;
    for (StepSlot &sl : e->slot) {
        for (EJob *j : sl.fin) delete j;
        if (sl.h_pkt) (void)hipHostFree(sl.h_pkt);
        if (sl.h_res) (void)hipHostFree(sl.h_res);
        if (sl.e0) (void)hipEventDestroy(sl.e0);
        if (sl.e1) (void)hipEventDestroy(sl.e1);
    }
    if (e->arena) (void)hipFree(e->arena);
    if (e->d_pkt) (void)hipFree(e->d_pkt);
    if (e->d_tp) (void)hipFree(e->d_tp);
    if (e->big) (void)hipFree(e->big);
    if (e->big_prefix) (void)hipFree(e->big_prefix);
    if (e->big_hdr) (void)hipFree(e->big_hdr);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->join) (void)hipEventDestroy(e->join);
    if (e->s2) (void)hipStreamDestroy(e->s2);
    if (e->s) (void)hipStreamDestroy(e->s);
    delete e;
}

int engine_active(const StepEngine *e) {
    return (int)(e->active.size() + e->incoming.size()) + (e->slot[0].busy ? 1 : 0) + (e->slot[1].busy ? 1 : 0);
}

/* Admission touches no device state: the slice is carved from the arena and
 * the query blob kept on the host until the next step's packet carries it. */
static bool engine_admit_hash(StepEngine *e, const DPInput &in, DPAlgo algo, uint64_t key) {
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    const bool rank = in.memo == GQ_MEMO_RANK;
    if (rank && !h.o_rank) return false;
    uint64_t total = 0, list = 0;
    for (int k = 1; k <= n; k++) {
        total += in.tree_counts[k];
        if (k >= 2) list += in.tree_counts[k];
    }
    if (2 * total > (1ull << 31)) return false;
    uint64_t cap;
    size_t o_blob, o_keys, o_vals, o_left, o_sets, o_slot;
    const size_t bytes = hash_slice_layout(total, list, h.words, rank, &cap, &o_blob, &o_keys, &o_vals, &o_left,
                                           &o_sets, &o_slot);
    size_t off;
    if (!e->alloc(bytes, &off)) return false;
    char *base = e->arena + off;
    EJob *j = new EJob();
    j->key = key;
    j->n = n;
    j->cur_k = 2;
    j->algo = algo;
    j->off = off;
    j->bytes = bytes;
    j->words = h.words;
    j->blob.assign(in.blob.begin(), in.blob.begin() + h.words);
    j->hash = true;
    j->expand = true;
    j->memo_kind = rank ? GQ_MEMO_RANK : GQ_MEMO_HASH;
    j->cap = rank ? 0 : cap;
    j->counts = in.tree_counts;
    BatchJob &J = j->J;
    memset(&J, 0, sizeof(J));
    J.h = h;
    J.blob = (const uint32_t *)(base + o_blob);
    J.keys = rank ? nullptr : (uint32_t *)(base + o_keys);
    J.hmask = rank ? 0u : (uint32_t)(cap - 1);
    J.vals = (float4 *)(base + o_vals);
    J.left = (uint32_t *)(base + o_left);
    J.lvl_sets = (uint32_t *)(base + o_sets);
    J.lvl_slot = (uint32_t *)(base + o_slot);
    J.counters = (uint32_t *)base;
    J.nodes = (DPNode *)(J.counters + 64);
    uint64_t acc = 0;
    for (int k = 0; k <= 32; k++) {
        J.lvl_off[k] = (uint32_t)acc;
        if (k >= 2 && k <= n) acc += in.tree_counts[k];
    }
    J.bicc = dp_mode(in, algo);
    J.tree_iters = tuning_get().tree_iters;
    J.cur_k = 2;
    e->incoming.push_back(j);
    return true;
}

bool engine_admit(StepEngine *e, const DPInput &in, DPAlgo algo, uint64_t key) {
    const BlobHdr &h = in.hdr;
    const int n = h.n_rels;
    if (!fused_eligible(in, algo) || (size_t)h.words * 4 > ENGINE_BLOB_MAX || in.blob.size() < (size_t)h.words)
        return false;
    const bool hash = (in.memo == GQ_MEMO_HASH || in.memo == GQ_MEMO_RANK) && h.is_tree &&
                      (int)in.tree_counts.size() == n + 1;
    if (hash) return engine_admit_hash(e, in, algo, key);
    size_t o_blob, o_vals, o_left, o_conn, o_sets, o_best, o_first;
    const size_t bytes = slice_layout(n, h.words, &o_blob, &o_vals, &o_left, &o_conn, &o_sets, &o_best, &o_first);
    size_t off;
    if (!e->alloc(bytes, &off)) return false;
    char *base = e->arena + off;
    EJob *j = new EJob();
    j->key = key;
    j->n = n;
    j->cur_k = 2;
    j->algo = algo;
    j->off = off;
    j->bytes = bytes;
    j->words = h.words;
    j->blob.assign(in.blob.begin(), in.blob.begin() + h.words);
    BatchJob &J = j->J;
    memset(&J, 0, sizeof(J));
    J.h = h;
    J.blob = (const uint32_t *)(base + o_blob);
    J.vals = (float4 *)(base + o_vals);
    J.left = (uint32_t *)(base + o_left);
    J.conn = (uint32_t *)(base + o_conn);
    J.conn_words = (uint32_t)std::max<uint64_t>(1, (1ull << n) / 32);
    J.lvl_sets = (uint32_t *)(base + o_sets);
    J.lvl_best = (unsigned long long *)(base + o_best);
    J.lvl_first = (uint32_t *)(base + o_first);
    J.counters = (uint32_t *)base;
    J.nodes = (DPNode *)(J.counters + 64);
    uint64_t acc = 0;
    for (int k = 0; k <= 32; k++) {
        J.lvl_off[k] = (uint32_t)acc;
        if (k >= 2 && k <= n) acc += binom64(n, k);
    }
    J.bicc = dp_mode(in, algo);
    J.tree_iters = tuning_get().tree_iters;
    J.cur_k = 2;
    j->hash = false;
    j->cap = 0;
    j->counts = in.tree_counts;
    /* a tree with the per-edge split table and exact level sizes builds each
     * level by expansion as it reaches it, instead of filtering all 2^n
     * integers when admitted (its evaluator never reads the connectivity
     * bitmap) */
    j->expand = fast_tree_job(h) && (int)in.tree_counts.size() == n + 1;
    J.expand = j->expand ? 1u : 0u;
    e->incoming.push_back(j);
    return true;
}

#define ECHK(x)                                                                            \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, #x); \
            return GQ_ERR_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

namespace {
size_t al64(size_t x) { return (x + 63) & ~(size_t)63; }

/* Enqueues one step: every active DP (the newly admitted ones last) advances
 * one level.  The host writes one packet into the slot's pinned buffer
 *   jobs[na] | expansion prefix[na+1] | tree task prefix[na+1] |
 *   filter chunk prefix[nn+1] | blob offsets[nn] | finishing job indexes[nf] |
 *   blobs of the new DPs
 * and the step starts with k_stage copying it to the device; the walk of
 * the finishing DPs writes their results straight into the slot.  The arena slices of the
 * finishing DPs are released at once: whatever the next steps place there
 * runs after this step's walk on the same stream. */
int engine_enqueue(StepEngine *e, int nn) {
    int rc;
    hipStream_t s = e->s;
    StepSlot &S = e->slot[e->cur];
    const int na = (int)e->active.size();
    const int j_new = na - nn;
    std::vector<uint32_t> fin_idx;
    size_t blob_words = 0;
    bool any_expand = false;
    for (int i = 0; i < na; i++) {
        const EJob *j = e->active[i];
        if (j->cur_k >= j->n) fin_idx.push_back((uint32_t)i);
        any_expand = any_expand || j->expand;
        if (i >= j_new) blob_words += j->blob.size();
    }
    const int nf = (int)fin_idx.size();
    const size_t o_exp = al64(sizeof(BatchJob) * (size_t)na);
    const size_t o_ttp = al64(o_exp + 4 * ((size_t)na + 1));
    const size_t o_chunk = al64(o_ttp + 4 * ((size_t)na + 1));
    const size_t o_bsrc = al64(o_chunk + 4 * ((size_t)nn + 1));
    const size_t o_fin = al64(o_bsrc + 4 * (size_t)nn);
    const size_t o_blob = al64(o_fin + 4 * (size_t)nf);
    const size_t pkt_bytes = (o_blob + 4 * blob_words + 15) & ~(size_t)15;
    const unsigned host_flags = hipHostMallocMapped | hipHostMallocCoherent;
    if ((rc = ensure_pinned(&S.h_pkt, S.cap_h_pkt, pkt_bytes, host_flags))) return rc;
    if ((rc = ensure_pinned(&S.h_res, S.cap_h_res, std::max<size_t>((size_t)nf, 1) * RES_WORDS, host_flags)))
        return rc;
    uint8_t *P = S.h_pkt;
    BatchJob *pj = (BatchJob *)P;
    uint32_t *pexp = (uint32_t *)(P + o_exp), *pttp = (uint32_t *)(P + o_ttp), *pchunk = (uint32_t *)(P + o_chunk),
             *pbsrc = (uint32_t *)(P + o_bsrc), *pfin = (uint32_t *)(P + o_fin), *pblob = (uint32_t *)(P + o_blob);
    size_t lds = 0, lds_tree = 0, big_lds_max = 0;
    uint64_t ub = 0, ub_tree = 0, big_cap = 0;
    pexp[0] = 0;
    pttp[0] = 0;
    for (int i = 0; i < na; i++) {
        EJob *j = e->active[i];
        j->J.cur_k = j->cur_k;
        pj[i] = j->J;
        const BlobHdr &h = j->J.h;
        const bool exact = (int)j->counts.size() == j->n + 1;
        const uint64_t lvl = exact ? j->counts[j->cur_k] : binom64(j->n, j->cur_k);
        uint64_t tt = 0;
        if (fast_tree_job(h)) {
            lds_tree = std::max(lds_tree, tree_lds(h.words, 1));
            tt = tree_tasks(lvl, j->cur_k, j->J.tree_iters);
        } else {
            lds = std::max(lds, std::max(eval_lds_bicc(h.words), eval_lds_dpsub(h.words, 0)));
            const uint64_t spt = task_sets_of(h.is_tree != 0, j->J.bicc != 0, j->cur_k, j->J.tree_iters);
            ub += (lvl + spt - 1) / spt;          /* tasks of a bound on the level size */
        }
        ub_tree += tt;
        if (ub_tree > 0x7fffffffull) {
            set_error("too many tasks in one step");
            return GQ_ERR_UNSUPPORTED;
        }
        pttp[i + 1] = (uint32_t)ub_tree;
        const uint64_t src = j->expand ? j->counts[j->cur_k - 1] : 0;
        pexp[i + 1] = pexp[i] + (uint32_t)((src + BLOCK - 1) / BLOCK);
        if (j->J.bicc == 1 && !h.is_tree && j->cur_k >= 15) {
            big_cap += binom64(j->n, j->cur_k);
            big_lds_max = std::max(big_lds_max, big_lds(h.words));
        }
    }
    if (big_cap > 0xffffffffull) {
        set_error("step too large for the big-set list");
        return GQ_ERR_UNSUPPORTED;
    }
    pchunk[0] = 0;
    size_t bw = 0;
    for (int i = 0; i < nn; i++) {
        const EJob *j = e->active[j_new + i];
        const uint64_t ns = 1ull << j->n;
        /* expansion jobs build their level lists step by step */
        pchunk[i + 1] = pchunk[i] + (j->expand ? 0u : (uint32_t)((ns + FILT_CHUNK - 1) / FILT_CHUNK));
        pbsrc[i] = (uint32_t)bw;
        memcpy(pblob + bw, j->blob.data(), 4 * j->blob.size());
        bw += j->blob.size();
    }
    for (int i = 0; i < nf; i++) pfin[i] = fin_idx[i];
    /* device buffers grow only with the previous step drained (a freed
     * buffer may still be in use by it) */
    const bool grow = e->cap_pkt < pkt_bytes || e->cap_tp < 2 * ((size_t)na + 1) ||
                      (big_cap && (e->cap_big < big_cap || e->cap_bigp < big_cap + 1));
    if (grow) {
        ECHK(hipStreamSynchronize(s));
        if ((rc = ensure(&e->d_pkt, e->cap_pkt, pkt_bytes))) return rc;
        if ((rc = ensure(&e->d_tp, e->cap_tp, 2 * ((size_t)na + 1)))) return rc;
        if (big_cap) {
            if ((rc = ensure(&e->big, e->cap_big, big_cap))) return rc;
            if ((rc = ensure(&e->big_prefix, e->cap_bigp, big_cap + 1))) return rc;
        }
    }
    uint8_t *D = e->d_pkt;
    const BatchJob *d_jobs = (const BatchJob *)D;
    ECHK(hipEventRecord(S.e0, s));
    {
        const uint32_t n16 = (uint32_t)(pkt_bytes / 16);
        hipLaunchKernelGGL(k_stage, dim3(std::max<uint32_t>(1, std::min<uint32_t>((n16 + BLOCK - 1) / BLOCK, 64))),
                           dim3(BLOCK), 0, s, (const uint4 *)P, (uint4 *)D, n16);
    }
    /* 1. DPs admitted since the last step: keys, base relations, filter */
    if (nn) {
        uint64_t keys = 0;
        for (int i = j_new; i < na; i++) keys += e->active[i]->cap;   /* 0 unless open addressing */
        if (keys)
            hipLaunchKernelGGL(k_clear_keys, dim3((uint32_t)std::min<uint64_t>((keys + BLOCK - 1) / BLOCK, 4096)),
                               dim3(BLOCK), 0, s, d_jobs + j_new, nn);
        hipLaunchKernelGGL(k_batch_init, dim3(std::min(nn, 1024)), dim3(64), 0, s, d_jobs + j_new, nn,
                           (const uint32_t *)(D + o_blob), (const uint32_t *)(D + o_bsrc));
        const uint32_t fgrid = std::min<uint32_t>(pchunk[nn], 4096);
        if (fgrid)
            hipLaunchKernelGGL(k_batch_filter, dim3(fgrid), dim3(BLOCK), 0, s, d_jobs + j_new, nn,
                               (const uint32_t *)(D + o_chunk));
    }
    /* 2. expansion jobs generate the level they evaluate now */
    if (any_expand)
        hipLaunchKernelGGL(k_expand, dim3(std::max<uint32_t>(1, std::min<uint32_t>(pexp[na], 4096))), dim3(BLOCK), 0,
                           s, d_jobs, na, 0, (const uint32_t *)(D + o_exp));
    /* 3. one level of every job: tree jobs on s (k_batch_eval_tree, task
     * prefix from the exact level sizes), the others on s2 (k_batch_eval,
     * task prefix from the device's level counters) */
    if (ub) {
        uint32_t *tp_gen = e->d_tp;
        hipLaunchKernelGGL(k_step_plan, dim3(1), dim3(1024), 0, s, d_jobs, na, tp_gen, JOBS_GENERAL);
        ECHK(hipEventRecord(e->fork, s));
        ECHK(hipStreamWaitEvent(e->s2, e->fork, 0));
        const uint32_t tpl = (lds + 4 * (size_t)(na + 1) <= 64 * 1024) ? (uint32_t)(lds / 4) : 0u;
        const size_t lt = tpl ? lds + 4 * (size_t)(na + 1) : lds;
        const uint32_t grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(ub, 1), tuning_get().gen_grid_max);
        hipLaunchKernelGGL(k_batch_eval, dim3(grid), dim3(BLOCK), lt, e->s2, d_jobs, na, 0, tp_gen, tpl,
                           big_cap ? e->big : nullptr, (uint32_t)big_cap, e->big_hdr);
        if (big_cap) {
            hipLaunchKernelGGL(k_big_scan, dim3(1), dim3(1024), 0, e->s2, e->big, e->big_hdr, (uint32_t)big_cap,
                               e->big_prefix);
            hipLaunchKernelGGL(k_big_eval, dim3(1024), dim3(BLOCK), big_lds_max, e->s2, d_jobs, e->big,
                               e->big_hdr, e->big_prefix);
        }
        ECHK(hipEventRecord(e->join, e->s2));
    }
    if (ub_tree)
        hipLaunchKernelGGL(k_batch_eval_tree, dim3((uint32_t)ub_tree), dim3(BLOCK), std::max<size_t>(lds_tree, 16),
                           s, d_jobs, na, 0, (const uint32_t *)(D + o_ttp));
    if (ub) ECHK(hipStreamWaitEvent(s, e->join, 0));
    /* 4. DPs past their last level: plan walk, one copy of the results */
    if (nf)
        hipLaunchKernelGGL(k_batch_tree, dim3(nf), dim3(WAVE), 0, s, d_jobs, nf, (const uint32_t *)(D + o_fin),
                           S.h_res);
    ECHK(hipGetLastError());
    ECHK(hipEventRecord(S.e1, s));
    std::vector<EJob *> keep;
    keep.reserve(na);
    for (int i = 0, f = 0; i < na; i++) {
        EJob *j = e->active[i];
        if (f < nf && fin_idx[f] == (uint32_t)i) {
            f++;
            e->release(j->off, j->bytes);
            S.fin.push_back(j);
        } else {
            j->cur_k++;
            keep.push_back(j);
        }
    }
    e->active.swap(keep);
    S.busy = true;
    e->cur ^= 1;
    return GQ_OK;
}

/* Waits for the step of slot `si` and turns its finished DPs into results. */
int engine_collect(StepEngine *e, int si, std::vector<EngineDone> &done) {
    StepSlot &S = e->slot[si];
    if (!S.busy) return GQ_OK;
    ECHK(hipEventSynchronize(S.e1));
    S.busy = false;
    float ms = 0.0f;
    ECHK(hipEventElapsedTime(&ms, S.e0, S.e1));
    for (size_t i = 0; i < S.fin.size(); i++) {
        EJob *j = S.fin[i];
        const uint32_t *cnt = S.h_res + i * RES_WORDS;
        EngineDone d;
        d.key = j->key;
        d.rc = GQ_OK;
        const int n = j->n;
        bool counts_ok = true;
        if (j->expand)
            for (int k = 2; k <= n; k++) counts_ok = counts_ok && cnt[k] == j->counts[k];
        if (cnt[49]) {
            d.rc = (cnt[49] & 1u) ? GQ_ERR_UNSUPPORTED : GQ_ERR_DEVICE;
        } else if (!counts_ok || cnt[48] != (uint32_t)(2 * n - 1)) {
            d.rc = GQ_ERR_DEVICE;
        } else {
            const DPNode *nodes = (const DPNode *)(cnt + 64);
            d.tree.assign(nodes, nodes + 2 * n - 1);
            for (int k = 2; k <= n; k++) d.st.sets += cnt[k];
            d.st.candidates = j->expand ? d.st.sets : (1ull << n) - (uint64_t)n - 1ull;
            d.st.pairs = (uint64_t)cnt[40] | ((uint64_t)cnt[41] << 32);
            if (j->J.h.is_tree && (int)j->counts.size() == n + 1) d.st.pairs = tree_pairs(j->counts);
            d.st.memo_bytes = j->bytes;
            d.st.memo_kind = j->memo_kind;
            d.st.gpu_ms = ms;
        }
        delete j;
        done.push_back(std::move(d));
    }
    S.fin.clear();
    return GQ_OK;
}
}  // namespace

/* Enqueues the next step (when any DP is active) and returns the DPs of the
 * step before it; with nothing to enqueue, drains the last step. */
int engine_step(StepEngine *e, std::vector<EngineDone> &done) {
    int rc;
    ECHK(hipSetDevice(e->device));
    const int nn = (int)e->incoming.size();
    for (EJob *j : e->incoming) e->active.push_back(j);
    e->incoming.clear();
    if (!e->active.empty()) {
        if ((rc = engine_enqueue(e, nn))) return rc;
        return engine_collect(e, e->cur, done);     /* the older step: slot cur after the flip */
    }
    if ((rc = engine_collect(e, e->cur ^ 1, done))) return rc;
    return engine_collect(e, e->cur, done);
}
}  // namespace gq
