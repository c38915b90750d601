/*
 * gq_wide.hip -- the wide DP: DPsub / MPDP over one tree query of 33 to 64
 * relations, relation i = bit i of a uint64_t set, on CDNA4 (gfx950, wave64).
 *
 * The reference's GPU DPsub takes such queries as Bitmapset64
 * (gpuqo_main_internal.cu:178-184): per level it unranks every k-subset of
 * the n relations (gpuqo_dpsub.cuh:95-126, C(n, k) of them), keeps the
 * connected ones (gpuqo_dpsub_filtered_kernels.cuh:30-109) and evaluates
 * them (:111-240) against an open-addressing memo
 * (gpuqo_hashtable.cuh:119-166).  Above 32 relations that pass over 2^n
 * candidates is what bounds it; here no candidate that is not a connected
 * set is ever formed:
 *
 *   levels   in a tree numbered breadth-first, the highest member v of a
 *            connected set T of two or more relations has its parent in T
 *            and is a leaf of T, so T - v is connected: T is built exactly
 *            once, from S = T - v, as S + v for a child v of a member of S
 *            with v above every member of S.  The host knows each level's
 *            size in advance (the subtree-count polynomial), so the lists
 *            and the memo are sized exactly, with no host synchronisation.
 *   memo     open addressing, 32-byte slots {key, left child, record},
 *            load at most 1/2, Fibonacci hashing and linear probing.  A
 *            set's slot is claimed (atomicCAS on the key) when the set is
 *            generated, one level ahead; lookups only ever look for keys of
 *            earlier launches, whose probe chains later inserts never
 *            shorten.
 *   kw_level one launch per level k: the connected sets of size k are
 *            evaluated -- k - 1 lanes per set, lane j takes the edge above
 *            the j-th non-top member c and costs both orientations of the
 *            split (S - sub(c), sub(c)) -- and, in the same launch, their
 *            extensions are inserted as level k + 1.  The lanes of a set
 *            reduce (first valid split, first minimum of (cost, rank)) with
 *            shuffles; ranks follow the reference's enumeration orders
 *            (below).
 *   kw_walk  one wave walks the memo breadth-first from the full set.
 *
 * Split orders (ties keep the first split found, gpuqo_cpu_common.cuh:
 * 189-211):
 *   MPDP   blocks in find_blocks emission order (gpuqo_cpu_dpsub_bicc.cu:
 *          65-169): in a tree every edge is a block, emitted in the
 *          post-order of its lower end c in the DFS from the lowest member
 *          of S with ascending neighbours; within the block (parent(c), c)
 *          lb = {parent(c)} first (:180-232).  Rank 2 post(c) + o, post(c) =
 *          |S & desc(c)| - 1 + |S & lglob(c)|.
 *   DPsub  left side ascending (gpuqo_cpu_dpsub.cu:32-79), the rank being
 *          the left side itself; the valid splits of a tree set are exactly
 *          its edge splits.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "gq_device.h"
#include "gq_host.h"
#include "gq_internal.h"
#include "gq_wide.h"

namespace gq {

#define WW 64                 /* wave */
#define WBLOCK 256            /* threads per workgroup of kw_level */
#define WGRID_MAX 8192u       /* workgroups of a kw_level launch (they loop over the level) */
#define WG_ECW 4              /* class-mask words of the general path (256 classes) */
#define WG_BLOCK_MAX 24       /* largest biconnected block that is not a simple cycle (2^m - 2 subsets) */

/* ------------------------------------------------------------ host side */

int build_wide_input(const HQuery<1> &q, WideInput &in) {
    const int n = q.n;
    if (n < 2 || n > GQ_WIDE_MAX_RELS) {
        set_error("the wide DP takes 2..%d relations (got %d)", GQ_WIDE_MAX_RELS, n);
        return GQ_ERR_UNSUPPORTED;
    }
    int edges = 0;
    for (int v = 0; v < n; v++) edges += __builtin_popcountll(q.adj[v].w[0]);
    /* a tree numbered parents first (gq_plan's BFS relabelling does that,
     * and so does the ascending renumbering of a subtree of such a tree in
     * IDP2) takes the tree tables; any other connected graph the general
     * path */
    bool tree = edges == 2 * (n - 1);
    std::vector<int> parent(n, -1);
    for (int v = 1; v < n && tree; v++) {
        const uint64_t lower = q.adj[v].w[0] & ((1ull << v) - 1ull);
        if (__builtin_popcountll(lower) != 1) tree = false;
        else parent[v] = __builtin_ctzll(lower);
    }
    HostView<1> hv;
    hv.build(q);
    const View<Set<1>> &v = hv.v;
    const int n_ec = v.n_ec, ecw = v.ecw, nm = (int)q.vars.size();
    if (ecw > (tree ? GQ_MAX_ECW : WG_ECW)) {
        set_error("the wide DP supports at most %d equivalence classes on %s (got %d)",
                  64 * (tree ? GQ_MAX_ECW : WG_ECW), tree ? "trees" : "graphs with cycles", n_ec);
        return GQ_ERR_UNSUPPORTED;
    }
    WideHdr &h = in.hdr;
    memset(&h, 0, sizeof(h));
    std::vector<uint32_t> &b = in.blob;
    b.clear();
    auto take = [&](int words) {
        const int o = (int)b.size();
        b.resize(b.size() + words, 0u);
        return o;
    };
    auto align = [&](int k) { while (b.size() % k) b.push_back(0u); };
    auto put_f = [&](int o, int i, float x) { memcpy(&b[o + i], &x, 4); };
    auto put_u64 = [&](int o, int i, uint64_t x) {
        b[o + 2 * i] = (uint32_t)x;
        b[o + 2 * i + 1] = (uint32_t)(x >> 32);
    };
    h.n_rels = n;
    h.n_ec = n_ec;
    h.n_mem = nm;
    h.ecw = ecw;
    h.p = q.p;
    h.composite = v.composite.w[0];
    h.o_rel_rows = take(64);
    h.o_rel_fksel = take(64);
    h.o_base = take(4 * 64);
    for (int i = 0; i < n; i++) {
        put_f(h.o_rel_rows, i, hv.rel_rows[i]);
        put_f(h.o_rel_fksel, i, hv.rel_fksel[i]);
        float st = q.startup[i], tot = q.total[i];
        base_cost(q.p, q.tuples[i], q.composite[i] != 0, &st, &tot);   /* cost_baserel */
        put_f(h.o_base, 4 * i, q.rows[i]);
        put_f(h.o_base, 4 * i + 1, st);
        put_f(h.o_base, 4 * i + 2, tot);
        b[h.o_base + 4 * i + 3] = (uint32_t)q.width[i];
    }
    align(2);
    h.o_inc = take(2 * n * ecw);
    for (int i = 0; i < n * ecw; i++) put_u64(h.o_inc, i, hv.inc[i]);
    h.o_ec_relids = take(2 * std::max(1, n_ec));
    h.o_ec_moff = take(std::max(1, n_ec));
    h.o_ec_soff = take(std::max(1, n_ec));
    h.o_ec_hasfk = take(std::max(1, (n_ec + 3) / 4));
    for (int e = 0; e < n_ec; e++) {
        put_u64(h.o_ec_relids, e, q.ec_relids[e].w[0]);
        b[h.o_ec_moff + e] = (uint32_t)hv.ec_moff[e];
        b[h.o_ec_soff + e] = (uint32_t)hv.ec_soff[e];
        ((uint8_t *)&b[h.o_ec_hasfk])[e] = hv.ec_hasfk[e];
    }
    align(2);
    h.o_mem_fks = take(2 * std::max(1, nm));
    h.o_mem_nd = take(std::max(1, nm));
    h.o_mem_inv_nd = take(std::max(1, nm));
    h.o_mem_skew = take(std::max(1, nm));
    h.o_mem_mcv = take(std::max(1, nm));
    h.o_mem_flags = take(std::max(1, (nm + 3) / 4));
    h.o_mem_idx_rows = take(std::max(1, nm));
    h.o_mem_idx_startup = take(std::max(1, nm));
    h.o_mem_idx_total = take(std::max(1, nm));
    for (int m = 0; m < nm; m++) {
        put_u64(h.o_mem_fks, m, q.fks[m].w[0]);
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
    h.pairs = v.pairs;
    align(4);
    h.o_pair_rec = take(h.pairs ? 12 * n_ec : 4);
    if (h.pairs) memcpy(&b[h.o_pair_rec], hv.pair_rec.data(), sizeof(PairRec) * 3 * (size_t)n_ec);
    h.o_adj = take(2 * 64);
    for (int u = 0; u < n; u++) put_u64(h.o_adj, u, q.adj[u].w[0]);
    h.is_tree = tree ? 1 : 0;
    if (!tree) {
        h.tree_fast = 0;
        align(4);
        h.words = (int)b.size();
        in.counts.clear();
        return GQ_OK;
    }
    /* tree tables: descendants, left-global sets, children */
    std::vector<uint64_t> desc(n, 0), lg(n, 0), kids(n, 0);
    for (int u = n - 1; u >= 0; u--) {
        desc[u] |= 1ull << u;
        if (u > 0) {
            desc[parent[u]] |= desc[u];
            kids[parent[u]] |= 1ull << u;
        }
    }
    for (int u = 1; u < n; u++) {              /* parents precede children */
        uint64_t l = lg[parent[u]];
        for (int s = 1; s < u; s++)
            if (parent[s] == parent[u]) l |= desc[s];
        lg[u] = l;
    }
    h.o_desc = take(2 * 64);
    h.o_lglob = take(2 * 64);
    h.o_kids = take(2 * 64);
    for (int u = 0; u < n; u++) {
        put_u64(h.o_desc, u, desc[u]);
        put_u64(h.o_lglob, u, lg[u]);
        put_u64(h.o_kids, u, kids[u]);
    }
    /* per-edge split tables when every class of two or more members is one
     * tree edge and no edge carries more than 4 classes: the crossing
     * classes of the split at edge (parent(c), c) are then exactly those on
     * that edge, so the selectivity (estimate_join_selectivity, the product
     * in ascending class order of ec_selectivity, which reads the class's
     * members on each side only) and the inner member slots depend on (c,
     * orientation) alone */
    std::vector<std::vector<int>> on_edge(n);
    bool fast = true;
    for (int e = 0; e < n_ec && fast; e++) {
        const uint64_t rel = q.ec_relids[e].w[0];
        const int pc = __builtin_popcountll(rel);
        if (pc < 2) continue;
        const int a = __builtin_ctzll(rel), c = 63 - __builtin_clzll(rel);
        if (pc != 2 || parent[c] != a) fast = false;
        else on_edge[c].push_back(e);
    }
    for (int c = 1; c < n && fast; c++)
        if (on_edge[c].size() > 4) fast = false;
    h.tree_fast = fast ? 1 : 0;
    h.o_tedge = 0;
    if (fast) {
        h.o_tedge = take(8 * 64);
        for (int c = 1; c < n; c++) {
            const Set<1> sc = Set<1>::nth(c), sp = Set<1>::nth(parent[c]);
            float sel0 = 1.0f, sel1 = 1.0f;
            uint16_t mi0[4] = {0, 0, 0, 0}, mi1[4] = {0, 0, 0, 0};
            int k = 0;
            for (int e : on_edge[c]) {                   /* ascending class order */
                sel0 *= ec_selectivity(v, e, sp, sc);
                sel1 *= ec_selectivity(v, e, sc, sp);
                const Set<1> ec = q.ec_relids[e];
                mi0[k] = (uint16_t)(hv.ec_moff[e] + rank_below(ec, sc));
                mi1[k] = (uint16_t)(hv.ec_moff[e] + rank_below(ec, sp));
                k++;
            }
            uint32_t *t = b.data() + h.o_tedge + 8 * c;
            memcpy(&t[0], &sel0, 4);
            memcpy(&t[1], &sel1, 4);
            t[2] = (uint32_t)k | ((uint32_t)k << 8);
            t[3] = 0;
            memcpy(&t[4], mi0, 8);
            memcpy(&t[6], mi1, 8);
        }
    }
    align(4);
    h.words = (int)b.size();
    /* connected sets per size: P_u(x) = x prod_{c child of u} (1 + P_c(x)),
     * summed over u (each subtree counted at its top vertex); every
     * coefficient counts sets of relations, so it fits 64 bits */
    std::vector<std::vector<uint64_t>> P(n);
    in.counts.assign(n + 1, 0ull);
    for (int u = n - 1; u >= 0; u--) {
        std::vector<uint64_t> poly(n + 1, 0ull);
        poly[1] = 1;
        for (uint64_t ch = kids[u]; ch; ch &= ch - 1) {
            const int c = __builtin_ctzll(ch);
            std::vector<uint64_t> r(n + 1, 0ull);
            for (int a = 1; a <= n; a++) {
                if (!poly[a]) continue;
                r[a] += poly[a];                          /* c not taken */
                for (int bb = 1; a + bb <= n; bb++) r[a + bb] += poly[a] * P[c][bb];
            }
            poly.swap(r);
        }
        P[u] = poly;
        for (int k = 0; k <= n; k++) in.counts[k] += poly[k];
    }
    return GQ_OK;
}

/* memo slots: a power of two at least twice the connected sets (and 256) */
static uint32_t wide_slots_log2(const WideInput &in) {
    uint64_t total = 0;
    for (uint64_t c : in.counts) total += c;
    uint32_t lg = 8;
    while (lg < 31 && (1ull << lg) < 2 * total) lg++;
    return lg;
}

uint64_t wide_dp_bytes(const WideInput &in) {
    uint64_t total = 0;
    for (uint64_t c : in.counts) total += c;
    const uint32_t lg = wide_slots_log2(in);
    return (32ull << lg) + 4ull * total;
}

/* ------------------------------------------------------------ device side */

/* position of the member of x that has j members below it (j < popcount(x)) */
__device__ __forceinline__ int select64(uint64_t x, uint32_t j) {
    int p = 0;
    uint32_t c = __builtin_popcount((uint32_t)x);
    if (j >= c) {
        j -= c;
        x >>= 32;
        p = 32;
    }
    uint32_t y = (uint32_t)x;
    c = __builtin_popcount(y & 0xffffu);
    if (j >= c) { j -= c; y >>= 16; p += 16; }
    c = __builtin_popcount(y & 0xffu);
    if (j >= c) { j -= c; y >>= 8; p += 8; }
    c = __builtin_popcount(y & 0xfu);
    if (j >= c) { j -= c; y >>= 4; p += 4; }
    c = __builtin_popcount(y & 0x3u);
    if (j >= c) { j -= c; y >>= 2; p += 2; }
    return (j >= (y & 1u)) ? p + 1 : p;
}

__device__ __forceinline__ unsigned long long shfl_down64(unsigned long long x, uint32_t o) {
    const uint32_t lo = __shfl_down((uint32_t)x, o), hi = __shfl_down((uint32_t)(x >> 32), o);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long shfl64(unsigned long long x, uint32_t src) {
    const uint32_t lo = __shfl((uint32_t)x, (int)src), hi = __shfl((uint32_t)(x >> 32), (int)src);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ View<uint64_t> make_wview(const WideHdr &h, const uint32_t *s) {
    View<uint64_t> v;
    v.n_rels = h.n_rels;
    v.n_ec = h.n_ec;
    v.ecw = h.ecw;
    v.p = h.p;
    v.rel_rows = (const float *)(s + h.o_rel_rows);
    v.rel_fksel = (const float *)(s + h.o_rel_fksel);
    v.composite = h.composite;
    v.inc = (const uint64_t *)(s + h.o_inc);
    v.ec_relids = (const uint64_t *)(s + h.o_ec_relids);
    v.ec_moff = (const int *)(s + h.o_ec_moff);
    v.ec_soff = (const int *)(s + h.o_ec_soff);
    v.ec_hasfk = (const uint8_t *)(s + h.o_ec_hasfk);
    v.mem_fks = (const uint64_t *)(s + h.o_mem_fks);
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

/* error bits of a wide DP (misc[WERR]) */
#define WERR_PROBE 1u      /* a lookup missed a set of an earlier level */
#define WERR_FULL 2u       /* an insert found no free slot, or a level list overflowed */
#define WERR_WALK 4u       /* the plan walk met an inconsistent entry */
#define WERR 66            /* misc[0..65]: sets generated per level; misc[66]: error bits */

/* Base relations (level 1) and their extensions (level 2). */
__global__ void __launch_bounds__(WW) kw_base(WideHdr h, const uint32_t *blob, WTab t, uint32_t *lst,
                                              uint32_t off2, uint32_t cap2, uint32_t *misc) {
    const int i = threadIdx.x;
    if (i >= h.n_rels) return;
    const uint64_t S = 1ull << i;
    const uint32_t sl = winsert(t, S);
    if (sl == ~0u) {
        atomicOr(&misc[WERR], WERR_FULL);
        return;
    }
    const float *bv = (const float *)(blob + h.o_base) + 4 * i;
    Rel r;
    r.rows = bv[0];
    r.startup = bv[1];
    r.total = bv[2];
    r.width = __float_as_int(bv[3]);
    r.hs = 0u;
    t.s[sl].rec = record_of(r, h.p.work_mem);
    t.s[sl].left = 0ull;
    lst[i] = sl;
    const uint64_t *kids = (const uint64_t *)(blob + h.o_kids);
    for (uint64_t c = kids[i]; c; c &= c - 1ull) {
        const uint32_t s2 = winsert(t, S | (c & (0ull - c)));
        const uint32_t pos = atomicAdd(&misc[2], 1u);
        if (s2 == ~0u || pos >= cap2) {
            atomicOr(&misc[WERR], WERR_FULL);
            continue;
        }
        lst[off2 + pos] = s2;
    }
}

/* Level k: evaluate its n_k sets (list at off_k) and, below the top level,
 * insert their extensions as level k + 1 (list at off_n, cap_n of them). */
template <bool DPSUB, bool FAST>
__global__ void __launch_bounds__(WBLOCK) kw_level(WideHdr h, const uint32_t *blob, WTab t, uint32_t *lst, int k,
                                                   uint32_t off_k, uint32_t n_k, uint32_t off_n, uint32_t cap_n,
                                                   uint32_t *misc) {
    extern __shared__ __align__(16) uint32_t wsm[];
    for (int i = threadIdx.x; i < h.words; i += blockDim.x) wsm[i] = blob[i];
    __syncthreads();
    const View<uint64_t> q = make_wview(h, wsm);
    const uint64_t *desc = (const uint64_t *)(wsm + h.o_desc);
    const uint64_t *lglob = (const uint64_t *)(wsm + h.o_lglob);
    const uint64_t *kids = (const uint64_t *)(wsm + h.o_kids);
    const uint32_t L = (uint32_t)(k - 1), spw = WW / L;
    const uint32_t lane = threadIdx.x & (WW - 1), seg = lane / L, j = lane - seg * L, lead = seg * L;
    const bool in_seg = seg < spw;
    const uint32_t wave = blockIdx.x * (WBLOCK / WW) + threadIdx.x / WW, nwaves = gridDim.x * (WBLOCK / WW);
    const bool expand = k < h.n_rels;
    const uint32_t NONE = 0xffffffffu;
    for (uint32_t base = wave * spw; base < n_k; base += nwaves * spw) {
        const uint32_t idx = base + seg;
        const bool live = in_seg && idx < n_k;
        uint32_t slot = 0;
        uint64_t S = 0, sub = 0, oth = 0;
        int c = 0;
        unsigned long long tie0 = ~0ull, tie1 = ~0ull;
        bool nan0 = false, nan1 = false;
        Rel r0, r1;
        if (live) {
            slot = lst[off_k + idx];
            S = t.s[slot].key;
            const uint64_t top = S & (0ull - S);
            c = select64(S ^ top, j);
            sub = S & desc[c];
            oth = S ^ sub;
            const uint32_t is = wfind(t, sub), io = wfind(t, oth);
            if (is == ~0u || io == ~0u) {
                atomicOr(&misc[WERR], WERR_PROBE);
            } else {
                const Rel Vs = rel_of(t.s[is].rec), Vo = rel_of(t.s[io].rec);
                if (FAST) {
                    const uint32_t *te = wsm + h.o_tedge + 8 * c;
                    EdgeProv p0, p1;
                    p0.sel = __uint_as_float(te[0]);
                    p1.sel = __uint_as_float(te[1]);
                    p0.nq = p1.nq = (int)(te[2] & 0xffu);
                    p0.n_mi = p1.n_mi = (int)((te[2] >> 8) & 0xffu);
                    p0.mi = (const uint16_t *)(te + 4);
                    p1.mi = (const uint16_t *)(te + 6);
                    r0 = join_prov(q, oth, Vo, sub, Vs, p0);      /* parent's side outer */
                    r1 = join_prov(q, sub, Vs, oth, Vo, p1);
                } else {
                    r0 = join<0>(q, oth, Vo, sub, Vs);
                    r1 = join<0>(q, sub, Vs, oth, Vo);
                }
                if (DPSUB) {
                    tie0 = oth;
                    tie1 = sub;
                } else {
                    const uint32_t post =
                        (uint32_t)(__builtin_popcountll(S & desc[c]) - 1 + __builtin_popcountll(S & lglob[c]));
                    tie0 = 2ull * post;
                    tie1 = 2ull * post + 1ull;
                }
                nan0 = r0.total != r0.total;
                nan1 = r1.total != r1.total;
            }
        }
        /* (a) the first split (lowest rank) with its NaN flag; (b) the first
         * minimum of (cost, rank) over non-NaN costs */
        unsigned long long tA = tie0 < tie1 ? tie0 : tie1;
        bool nA = tie0 < tie1 ? nan0 : nan1;
        uint32_t oB = NONE;
        unsigned long long tB = ~0ull;
        if (tie0 != ~0ull && !nan0) { oB = ord_key(r0.total); tB = tie0; }
        if (tie1 != ~0ull && !nan1) {
            const uint32_t o1 = ord_key(r1.total);
            if (o1 < oB || (o1 == oB && tie1 < tB)) { oB = o1; tB = tie1; }
        }
        for (uint32_t o = 1; o < L; o <<= 1) {
            const unsigned long long xtA = shfl_down64(tA, o), xtB = shfl_down64(tB, o);
            const bool xnA = __shfl_down((int)nA, o) != 0;
            const uint32_t xoB = __shfl_down(oB, o);
            if (j + o < L) {
                if (xtA < tA) { tA = xtA; nA = xnA; }
                if (xoB < oB || (xoB == oB && xtB < tB)) { oB = xoB; tB = xtB; }
            }
        }
        const unsigned long long win0 = (nA || oB == NONE) ? tA : tB;
        const unsigned long long win = shfl64(win0, lead);
        if (live && win != ~0ull) {
            if (tie0 == win) {
                t.s[slot].rec = record_of(r0, h.p.work_mem);
                t.s[slot].left = oth;
            } else if (tie1 == win) {
                t.s[slot].rec = record_of(r1, h.p.work_mem);
                t.s[slot].left = sub;
            }
        }
        if (!expand) continue;
        /* extensions: children of members of S above its highest member */
        unsigned long long acc = 0ull;
        if (live) {
            acc = kids[c];
            if (j == 0) acc |= kids[__builtin_ctzll(S)];
        }
        for (uint32_t o = 1; o < L; o <<= 1) {
            const unsigned long long x = shfl_down64(acc, o);
            if (j + o < L) acc |= x;
        }
        acc = shfl64(acc, lead);
        uint64_t C = 0ull;
        if (live) {
            const int mx = 63 - __builtin_clzll(S);
            C = mx >= 63 ? 0ull : (acc & (~0ull << (mx + 1)));
        }
        const uint32_t nc = (uint32_t)__builtin_popcountll(C);
        const uint32_t mine = (live && j < nc) ? (nc - j + L - 1u) / L : 0u;
        uint32_t incl = mine;                                  /* wave-inclusive scan */
        for (uint32_t o = 1; o < WW; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const uint32_t total = __shfl(incl, WW - 1);
        uint32_t first = 0;
        if (lane == WW - 1 && total) first = atomicAdd(&misc[k + 1], total);
        first = __shfl(first, WW - 1);
        uint32_t pos = first + incl - mine;
        for (uint32_t tt = j; tt < nc && live; tt += L, pos++) {
            const uint64_t T = S | (1ull << select64(C, tt));
            const uint32_t s2 = winsert(t, T);
            if (s2 == ~0u || pos >= cap_n) {
                atomicOr(&misc[WERR], WERR_FULL);
                continue;
            }
            lst[off_n + pos] = s2;
        }
    }
}

/* The plan, breadth-first from the set `root` (the full set, or a
 * connected sub-query of a memo snapshot): node i's children are allocated
 * after the nodes of its frontier. */
__global__ void __launch_bounds__(WW) kw_walk(WTab t, uint64_t root, WideNode *out, uint32_t *misc) {
    __shared__ unsigned long long ns[128];
    const uint32_t lane = threadIdx.x;
    if (lane == 0) ns[0] = root;
    __syncthreads();
    uint32_t fb = 0, fe = 1;
    while (fb < fe) {
        if (fe - fb > WW) {
            if (lane == 0) atomicOr(&misc[WERR], WERR_WALK);
            break;
        }
        const uint32_t i = fb + lane;
        const bool live = i < fe;
        uint64_t S = 0, Lset = 0;
        Rel r;
        r.rows = r.startup = r.total = 0.0f;
        r.width = 0;
        bool internal = false;
        if (live) {
            S = ns[i];
            const uint32_t sl = wfind(t, S);
            if (sl == ~0u) {
                atomicOr(&misc[WERR], WERR_WALK);
            } else {
                r = rel_of(t.s[sl].rec);
                Lset = t.s[sl].left;
                internal = __builtin_popcountll(S) > 1;
                if (internal && (Lset == 0ull || (Lset & ~S) != 0ull || Lset == S)) {
                    atomicOr(&misc[WERR], WERR_WALK);
                    internal = false;
                }
            }
        }
        const unsigned long long bal = __ballot(internal);
        const uint32_t ci = fe + 2u * (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        const uint32_t fe2 = fe + 2u * (uint32_t)__popcll(bal);
        if (fe2 > 128u) {
            if (lane == 0) atomicOr(&misc[WERR], WERR_WALK);
            break;
        }
        if (live) {
            WideNode nd;
            nd.set = S;
            nd.rows = r.rows;
            nd.startup = r.startup;
            nd.total = r.total;
            nd.width = r.width;
            nd.left = internal ? (int32_t)ci : -1;
            nd.right = internal ? (int32_t)ci + 1 : -1;
            out[i] = nd;
            if (internal) {
                ns[ci] = Lset;
                ns[ci + 1] = S ^ Lset;
            }
        }
        __syncthreads();
        fb = fe;
        fe = fe2;
    }
    if (lane == 0) misc[WERR + 1] = fe;
}

/* ------------------------------------------------- general join graphs
 *
 * Sets of a graph with cycles: a set T of k + 1 relations is an extension
 * S + v of each of its connected k-subsets S, so it is generated once per
 * such S; the first to claim its memo key appends it to level k + 1
 * (winsert_fresh).  Each set is evaluated by one wave:
 *   1. lane 0 splits it into biconnected blocks in find_blocks emission
 *      order (gpuqo_cpu_dpsub_bicc.cu:65-169; dfs_blocks64);
 *   2. the lanes take the blocks' splits strided: a simple cycle's arcs, or
 *      every non-empty proper subset lb of another block whose two parts
 *      are connected inside it (valid exactly then: a path between two
 *      members of a block never leaves it), with both sides grown by what
 *      hangs off their members (block_side64) -- the reference's
 *      grow(lowbit(lb), S - rb) / grow(lowbit(rb), S - lb) (:180-232);
 *   3. the crossing classes of a split are the classes on both sides, from
 *      per-member and per-block incidence masks the lanes build first;
 *   4. a butterfly reduction keeps (first valid split, with its NaN flag)
 *      and (first minimum of cost, rank) -- rank the split's place in the
 *      block order (MPDP) or its left side (DPsub, whose splits come in
 *      ascending left side, gpuqo_cpu_dpsub.cu:32-79; a valid split of
 *      DPsub has both sides connected, so it is one of these block
 *      splits) -- and the lane holding the winner writes the memo.
 * Blocks that are not simple cycles take at most WG_BLOCK_MAX relations
 * (WERR_BIG otherwise). */
#define WERR_BIG 8u        /* a block above WG_BLOCK_MAX that is not a cycle, or MPDP ranks past 64 bits */
#define WPAIRS 68          /* misc[68..69]: valid splits costed (u64) */

struct alignas(16) WgScratch {           /* one wave's set, in LDS */
    unsigned long long blocks[64], own[64], subb[65];
    unsigned long long incown[64 * WG_ECW], incout[64 * WG_ECW];
    unsigned long long cyc;
    uint32_t dfs[64];
    int nblk;
};

__device__ __forceinline__ uint64_t grow64(uint64_t from, uint64_t subset, const uint64_t *adj) {
    uint64_t V = 0, N = from;
    while (N) {
        const int i = __builtin_ctzll(N);
        V |= 1ull << i;
        N = (N | adj[i]) & subset & ~V;
    }
    return V;
}

__device__ __forceinline__ uint64_t pdep64(uint64_t j, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t bit = 1; mask; mask &= mask - 1ull, bit <<= 1)
        if (j & bit) r |= mask & (0ull - mask);
    return r;
}

__device__ __forceinline__ uint64_t pext64(uint64_t x, uint64_t mask) {
    uint64_t r = 0, bit = 1;
    for (; mask; mask &= mask - 1ull, bit <<= 1)
        if (x & mask & (0ull - mask)) r |= bit;
    return r;
}

/* Biconnected blocks of connected S: the DFS from S's lowest member with
 * ascending neighbours; a block is emitted when the tree edge (p, u)
 * finishes with low[u] >= disc[p] -- find_blocks' order.  blocks[b] = the
 * block, subb[b] = the vertices below its root p (its members but p and
 * what hangs off them), own[v] = v and what hangs below it outside v's own
 * block towards the root (own[root] = S); bit b of cyc: block b is a
 * simple cycle.  dfs[v] packs discovery time, low and DFS parent (8 bits
 * each); the vertex stack runs down from subb[64] (stack depth plus blocks
 * emitted stays within |S| <= 64). */
__device__ int dfs_blocks64(uint64_t S, const uint64_t *adj, WgScratch &W) {
#define STK(i) W.subb[64 - (i)]
    int nblk = 0;
    const int root = __builtin_ctzll(S);
    for (uint64_t x = S; x; x &= x - 1ull) W.own[__builtin_ctzll(x)] = 0ull;
    uint64_t visited = 1ull << root;
    uint32_t t = 1;
    int sp = 0;
    STK(sp++) = (unsigned long long)root;
    uint32_t du = 0u | (255u << 8) | (255u << 16);
    W.dfs[root] = du;
    uint64_t au = adj[root] & S;
    int u = root, cursor = -1;
    while (true) {
        const uint64_t gt = cursor < 0 ? ~0ull : (cursor >= 63 ? 0ull : (~0ull << (cursor + 1)));
        const uint64_t rem = au & gt;
        if (rem) {
            const int v = __builtin_ctzll(rem);
            if (!((visited >> v) & 1ull)) {
                du = t | (255u << 8) | ((uint32_t)u << 16);
                W.dfs[v] = du;
                t++;
                visited |= 1ull << v;
                STK(sp++) = (unsigned long long)v;
                au = adj[v] & S;
                u = v;
                cursor = -1;
            } else {
                if ((int)(du >> 16) != v) {       /* not the tree edge to u's parent */
                    const uint32_t dv = W.dfs[v];
                    if ((dv & 0xffu) < (du & 0xffu)) {
                        const uint32_t low = (du >> 8) & 0xffu;
                        const uint32_t nl = (dv & 0xffu) < low ? (dv & 0xffu) : low;
                        du = (du & 0xffff00ffu) | (nl << 8);
                        W.dfs[u] = du;
                    }
                }
                cursor = v;
            }
            continue;
        }
        if (u == root) break;
        const int p = (int)(du >> 16);
        uint32_t dp = W.dfs[p];
        const uint64_t ap = adj[p] & S;
        const uint32_t lowu = (du >> 8) & 0xffu;
        if (lowu >= (dp & 0xffu)) {
            uint64_t rest = 0, sub = 0;
            int v;
            do {                                   /* pop down to u */
                v = (int)STK(--sp);
                const uint64_t o = W.own[v] | (1ull << v);
                W.own[v] = o;
                rest |= 1ull << v;
                sub |= o;
            } while (v != u);
            W.blocks[nblk] = rest | (1ull << p);
            W.subb[nblk] = sub;
            nblk++;
            W.own[p] |= sub;
        }
        const uint32_t lowp = (dp >> 8) & 0xffu;
        if (lowu < lowp) {
            dp = (dp & 0xffff00ffu) | (lowu << 8);
            W.dfs[p] = dp;
        }
        cursor = u;
        u = p;
        du = dp;
        au = ap;
    }
#undef STK
    W.own[root] = S;
    uint64_t cyc = 0;
    for (int b = 0; b < nblk; b++) {
        const uint64_t blk = W.blocks[b];
        bool c = __builtin_popcountll(blk) >= 3;
        for (uint64_t x = blk; c && x; x &= x - 1ull) c = __builtin_popcountll(adj[__builtin_ctzll(x)] & blk) == 2;
        if (c) cyc |= 1ull << b;
    }
    W.cyc = cyc;
    return nblk;
}

/* grow(lowbit(x), S - (blk - x)) for a part x of block blk (below-root set
 * sb) that is connected inside the block */
__device__ __forceinline__ uint64_t block_side64(uint64_t S, uint64_t x, uint64_t sb, const unsigned long long *own) {
    uint64_t r = (x & ~sb) ? (S & ~sb) : 0ull;
    for (uint64_t y = x & sb; y; y &= y - 1ull) r |= own[__builtin_ctzll(y)];
    return r;
}

/* classes touching side x of block b (mask words m[0..ecw)) */
__device__ __forceinline__ void side_classes(const WgScratch &W, int b, uint64_t x, uint64_t sb, int ecw,
                                             uint64_t *m) {
    for (int w = 0; w < ecw; w++) m[w] = (x & ~sb) ? W.incout[b * WG_ECW + w] : 0ull;
    for (uint64_t y = x & sb; y; y &= y - 1ull) {
        const int v = __builtin_ctzll(y);
        for (int w = 0; w < ecw; w++) m[w] |= W.incown[v * WG_ECW + w];
    }
}

__device__ __forceinline__ unsigned long long xor64(unsigned long long x, int o) {
    const uint32_t lo = __shfl_xor((uint32_t)x, o), hi = __shfl_xor((uint32_t)(x >> 32), o);
    return ((unsigned long long)hi << 32) | lo;
}

/* Base relations (level 1, list at 0) and the edges (level 2, list at n). */
__global__ void __launch_bounds__(WW) kw_gbase(WideHdr h, const uint32_t *blob, WTab t, uint32_t *lst,
                                               uint32_t cap2, uint32_t *misc) {
    const int i = threadIdx.x;
    if (i >= h.n_rels) return;
    const uint64_t S = 1ull << i;
    bool fresh;
    const uint32_t sl = winsert_fresh(t, S, &fresh);
    if (sl == ~0u) {
        atomicOr(&misc[WERR], WERR_FULL);
        return;
    }
    const float *bv = (const float *)(blob + h.o_base) + 4 * i;
    Rel r;
    r.rows = bv[0];
    r.startup = bv[1];
    r.total = bv[2];
    r.width = __float_as_int(bv[3]);
    r.hs = 0u;
    t.s[sl].rec = record_of(r, h.p.work_mem);
    t.s[sl].left = 0ull;
    lst[i] = sl;
    const uint64_t *adj = (const uint64_t *)(blob + h.o_adj);
    const uint64_t up = i >= 63 ? 0ull : (adj[i] & (~0ull << (i + 1)));
    for (uint64_t c = up; c; c &= c - 1ull) {
        const uint32_t s2 = winsert_fresh(t, S | (c & (0ull - c)), &fresh);
        const uint32_t pos = atomicAdd(&misc[2], 1u);
        if (s2 == ~0u || pos >= cap2) {
            atomicOr(&misc[WERR], WERR_FULL);
            continue;
        }
        lst[h.n_rels + pos] = s2;
    }
}

/* Level k of a general graph: one wave per set (see above); below the top
 * level the set's extensions not yet in the memo become level k + 1 (list
 * at off_n, room for cap_n). */
template <bool DPSUB>
__global__ void __launch_bounds__(WBLOCK) kw_glevel(WideHdr h, const uint32_t *blob, WTab t, uint32_t *lst, int k,
                                                    uint32_t off_k, uint32_t n_k, uint32_t off_n, uint32_t cap_n,
                                                    uint32_t *misc) {
    extern __shared__ __align__(16) uint32_t wsm[];
    for (int i = threadIdx.x; i < h.words; i += blockDim.x) wsm[i] = blob[i];
    __syncthreads();
    const View<uint64_t> q = make_wview(h, wsm);
    const uint64_t *adj = (const uint64_t *)(wsm + h.o_adj);
    const uint64_t *inc = q.inc;
    const int ecw = h.ecw;
    WgScratch &W = ((WgScratch *)(wsm + ((h.words + 3) & ~3)))[threadIdx.x / WW];
    const uint32_t lane = threadIdx.x & (WW - 1);
    const uint32_t wave = blockIdx.x * (WBLOCK / WW) + threadIdx.x / WW, nwaves = gridDim.x * (WBLOCK / WW);
    const bool expand = k < h.n_rels;
    const uint32_t NONE = 0xffffffffu;
    for (uint32_t idx = wave; idx < n_k; idx += nwaves) {
        const uint32_t slot = lst[off_k + idx];
        const uint64_t S = t.s[slot].key;
        if (lane == 0) W.nblk = dfs_blocks64(S, adj, W);
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const int nblk = W.nblk;
        /* incidence of what each member owns and of each block's outside */
        if ((S >> lane) & 1ull) {
            for (int w = 0; w < ecw; w++) {
                uint64_t x = 0;
                for (uint64_t y = W.own[lane]; y; y &= y - 1ull) x |= inc[__builtin_ctzll(y) * ecw + w];
                W.incown[lane * WG_ECW + w] = x;
            }
        }
        if ((int)lane < nblk) {
            const uint64_t out = S & ~W.subb[lane];
            for (int w = 0; w < ecw; w++) {
                uint64_t x = 0;
                for (uint64_t y = out; y; y &= y - 1ull) x |= inc[__builtin_ctzll(y) * ecw + w];
                W.incout[lane * WG_ECW + w] = x;
            }
        }
        __builtin_amdgcn_wave_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        /* the set's enumerated splits and the rank room of the blocks */
        unsigned long long total = 0, rsum = 0;
        bool big = false;
        for (int b = 0; b < nblk; b++) {
            const int m = __builtin_popcountll(W.blocks[b]);
            const bool cy = (W.cyc >> b) & 1ull;
            if (!cy && m > WG_BLOCK_MAX) big = true;
            else {
                total += cy ? (unsigned long long)m * (m - 1) : (1ull << m) - 2ull;
                /* MPDP ranks count 2^m - 2 per block: a block of 64 (only a
                 * cycle through every relation) or a sum past 64 bits has
                 * no rank room; DPsub ranks by the left side instead */
                if (!DPSUB) {
                    const unsigned long long f = m >= 64 ? ~0ull : (1ull << m) - 2ull;
                    if (m >= 64 || rsum + f < rsum) big = true;
                    rsum += f;
                }
            }
        }
        if (big) {
            if (lane == 0) atomicOr(&misc[WERR], WERR_BIG);
            continue;
        }
        /* per lane: first valid split (rank, NaN flag, record, left) and the
         * first minimum of (cost, rank) over non-NaN costs */
        unsigned long long fr = ~0ull, mr = ~0ull, fl = 0, ml = 0;
        bool fnan = false;
        uint32_t mo = NONE, npairs = 0;
        Rel frec, mrec;
        int cb = 0;
        unsigned long long acc_i = 0, acc_r = 0;
        for (unsigned long long it = lane; it < total; it += WW) {
            uint64_t blk = W.blocks[cb];
            int m = __builtin_popcountll(blk);
            bool cy = (W.cyc >> cb) & 1ull;
            unsigned long long items = cy ? (unsigned long long)m * (m - 1) : (1ull << m) - 2ull;
            while (it >= acc_i + items) {               /* lanes move through the blocks in order */
                acc_i += items;
                acc_r += m >= 64 ? 0ull : (1ull << m) - 2ull;
                cb++;
                blk = W.blocks[cb];
                m = __builtin_popcountll(blk);
                cy = (W.cyc >> cb) & 1ull;
                items = cy ? (unsigned long long)m * (m - 1) : (1ull << m) - 2ull;
            }
            const unsigned long long local = it - acc_i;
            uint64_t lb;
            unsigned long long rank;
            if (cy) {        /* arc (start, len) from the lowest member towards its lower neighbour */
                const uint32_t m1 = (uint32_t)m - 1u;
                const uint32_t start = (uint32_t)(local / m1), len = (uint32_t)(local % m1) + 1u;
                uint64_t cur = blk & (0ull - blk), prev = 0;
                lb = 0;
                for (uint32_t st = 0; st < start + len; st++) {
                    if (st >= start) lb |= cur;
                    const uint64_t nb = adj[__builtin_ctzll(cur)] & blk & ~prev;
                    prev = cur;
                    cur = nb & (0ull - nb);
                }
                rank = acc_r + pext64(lb, blk) - 1ull;
            } else {
                lb = pdep64(local + 1ull, blk);
                rank = acc_r + local;
            }
            if (m >= 64) rank = 0;                      /* DPsub only: ranked by L below */
            const uint64_t rb = blk ^ lb;
            if (!cy && m > 2 &&
                (grow64(lb & (0ull - lb), lb, adj) != lb || grow64(rb & (0ull - rb), rb, adj) != rb))
                continue;
            const uint64_t sb = W.subb[cb];
            const uint64_t L = block_side64(S, lb, sb, W.own), R = block_side64(S, rb, sb, W.own);
            if (DPSUB) rank = L;
            const uint32_t il = wfind(t, L), ir = wfind(t, R);
            if (il == ~0u || ir == ~0u) {
                atomicOr(&misc[WERR], WERR_PROBE);
                continue;
            }
            uint64_t mL[WG_ECW], mR[WG_ECW], mm[GQ_MAX_ECW];
            side_classes(W, cb, lb, sb, ecw, mL);
            side_classes(W, cb, rb, sb, ecw, mR);
            int nq = 0;
            for (int w = 0; w < ecw; w++) {
                mm[w] = mL[w] & mR[w];
                nq += __builtin_popcountll(mm[w]);
            }
            const Rel r = join_mask<0>(q, L, rel_of(t.s[il].rec), R, rel_of(t.s[ir].rec), mm, nq);
            npairs++;
            const bool nan = r.total != r.total;
            if (rank < fr) {
                fr = rank;
                fnan = nan;
                frec = r;
                fl = L;
            }
            if (!nan) {
                const uint32_t o = ord_key(r.total);
                if (o < mo || (o == mo && rank < mr)) {
                    mo = o;
                    mr = rank;
                    mrec = r;
                    ml = L;
                }
            }
        }
        /* butterfly: every lane ends with the set's two winners */
        unsigned long long rfr = fr, rmr = mr;
        bool rfnan = fnan;
        uint32_t rmo = mo, rp = npairs;
        for (int o = 1; o < WW; o <<= 1) {
            const unsigned long long xfr = xor64(rfr, o), xmr = xor64(rmr, o);
            const bool xfn = __shfl_xor((int)rfnan, o) != 0;
            const uint32_t xmo = __shfl_xor(rmo, o);
            rp += __shfl_xor(rp, o);
            if (xfr < rfr) {
                rfr = xfr;
                rfnan = xfn;
            }
            if (xmo < rmo || (xmo == rmo && xmr < rmr)) {
                rmo = xmo;
                rmr = xmr;
            }
        }
        const bool nanfirst = rfnan || rmo == NONE;
        if (nanfirst ? (rfr != ~0ull && fr == rfr) : (mr == rmr)) {
            t.s[slot].rec = record_of(nanfirst ? frec : mrec, h.p.work_mem);
            t.s[slot].left = nanfirst ? fl : ml;
        }
        if (lane == 0 && rp) atomicAdd((unsigned long long *)(misc + WPAIRS), (unsigned long long)rp);
        if (!expand) continue;
        /* extensions S + v, v a neighbour of S: the first claim of a key
         * appends the set to level k + 1 */
        uint64_t nb = ((S >> lane) & 1ull) ? adj[lane] : 0ull;
        for (int o = 1; o < WW; o <<= 1) nb |= xor64(nb, o);
        const uint64_t C = nb & ~S;
        const uint32_t nc = (uint32_t)__builtin_popcountll(C);
        for (uint32_t base = 0; base < nc; base += WW) {
            bool fresh = false;
            uint32_t s2 = ~0u;
            if (base + lane < nc) {
                s2 = winsert_fresh(t, S | (1ull << select64(C, base + lane)), &fresh);
                if (s2 == ~0u) atomicOr(&misc[WERR], WERR_FULL);
            }
            const unsigned long long bal = __ballot(fresh);
            uint32_t first = 0;
            if (lane == 0 && bal) first = atomicAdd(&misc[k + 1], (uint32_t)__popcll(bal));
            first = __shfl(first, 0);
            if (fresh) {
                const uint32_t pos = first + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                if (pos < cap_n) lst[off_n + pos] = s2;
                else atomicOr(&misc[WERR], WERR_FULL);
            }
        }
    }
}

/* ------------------------------------------------------------ runner */

struct WideBufs {
    int device = -1;
    WSlot *tab = nullptr;
    uint32_t *lst = nullptr, *blob = nullptr, *misc = nullptr;
    WideNode *nodes = nullptr;
    size_t cap_tab = 0, cap_lst = 0, cap_blob = 0;
    uint32_t *h_misc = nullptr;
    WideNode *h_nodes = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
};
static WideBufs g_wide;      /* used under the device context's lock only */

#define WCHK(x)                                                                            \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, #x); \
            return GQ_ERR_DEVICE;                                                          \
        }                                                                                  \
    } while (0)

template <class T>
static int wensure(T **p, size_t &cap, size_t need) {
    if (*p && cap >= need) return GQ_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    const size_t bytes = std::max<size_t>(need, 1) * sizeof(T);
    if (hipMalloc((void **)p, bytes) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes failed", bytes);
        return GQ_ERR_NOMEM;
    }
    cap = std::max<size_t>(need, 1);
    return GQ_OK;
}

static int wide_init(WideBufs &w, int device) {
    if (w.device == device) return GQ_OK;
    if (w.device >= 0) {
        set_error("the wide DP runs on one device per process (device %d, asked for %d)", w.device, device);
        return GQ_ERR_UNSUPPORTED;
    }
    WCHK(hipMalloc((void **)&w.misc, 72 * 4));
    WCHK(hipMalloc((void **)&w.nodes, 128 * sizeof(WideNode)));
    WCHK(hipHostMalloc((void **)&w.h_misc, 72 * 4, hipHostMallocDefault));
    WCHK(hipHostMalloc((void **)&w.h_nodes, 128 * sizeof(WideNode), hipHostMallocDefault));
    WCHK(hipEventCreate(&w.e0));
    WCHK(hipEventCreate(&w.e1));
    w.device = device;
    return GQ_OK;
}

/* ------------------------------------------------------------ snapshots */

thread_local SnapReq *g_snap_req = nullptr;

static uint64_t full_set(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

uint32_t snap_lg(uint64_t n_sets) {
    uint32_t lg = 6;
    while (lg < 31 && (1ull << lg) < 2 * n_sets) lg++;
    return lg;
}

struct Perm64 {
    int8_t p[64];
};

__device__ __forceinline__ uint64_t perm64(uint64_t x, const Perm64 &pm) {
    uint64_t r = 0;
    for (; x; x &= x - 1ull) r |= 1ull << pm.p[__builtin_ctzll(x)];
    return r;
}

/* A wide DP's memo into a snapshot table, keys and left sides renumbered. */
__global__ void __launch_bounds__(WBLOCK) ks_remap(WTab from, WTab to, Perm64 pm, uint32_t *misc) {
    const uint64_t n = (uint64_t)from.mask + 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * WBLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WBLOCK) {
        const unsigned long long key = from.s[i].key;
        if (!key) continue;
        const uint32_t sl = winsert(to, perm64(key, pm));
        if (sl == ~0u) {
            atomicOr(&misc[WERR], WERR_FULL);
            continue;
        }
        to.s[sl].rec = from.s[i].rec;
        to.s[sl].left = perm64(from.s[i].left, pm);
        atomicAdd(&misc[0], 1u);
    }
}

/* Records into an empty table (an imported snapshot). */
__global__ void __launch_bounds__(WBLOCK) ks_fill(WTab to, const WSlot *recs, uint64_t n, uint32_t *misc) {
    for (uint64_t i = (uint64_t)blockIdx.x * WBLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WBLOCK) {
        bool fresh;
        const uint32_t sl = winsert_fresh(to, recs[i].key, &fresh);
        if (sl == ~0u || !fresh) {
            atomicOr(&misc[WERR], sl == ~0u ? WERR_FULL : WERR_PROBE);
            continue;
        }
        to.s[sl].rec = recs[i].rec;
        to.s[sl].left = recs[i].left;
    }
}

static int snap_alloc(SnapTable &t, uint64_t n_sets, int n_rels, int device, hipStream_t s) {
    t.lg = snap_lg(n_sets);
    t.n_rels = n_rels;
    t.device = device;
    t.n_sets = 0;
    const size_t bytes = sizeof(WSlot) << t.lg;
    if (hipMalloc(&t.tab, bytes) != hipSuccess) {
        t.tab = nullptr;
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes for the memo snapshot failed", bytes);
        return GQ_ERR_NOMEM;
    }
    if (hipMemsetAsync(t.tab, 0, bytes, s) != hipSuccess) {
        set_error("memo snapshot: clearing its table failed");
        return GQ_ERR_DEVICE;
    }
    return GQ_OK;
}

void snapshot_free(SnapTable &t) {
    if (t.tab) (void)hipFree(t.tab);
    t.tab = nullptr;
    t.n_sets = 0;
}

static WTab snap_wtab(const SnapTable &t) {
    WTab w;
    w.s = (WSlot *)t.tab;
    w.lg = t.lg;
    w.mask = (uint32_t)((1ull << t.lg) - 1ull);
    return w;
}

/* The wide DP's memo (total sets) into g_snap_req's table; called on the
 * DP's stream with the device context held. */
static int wide_snapshot(const WTab &memo, uint64_t total, int n, int device, hipStream_t s) {
    SnapReq &r = *g_snap_req;
    Perm64 pm;
    memcpy(pm.p, r.perm, 64);
    int rc = snap_alloc(r.out, total, n, device, s);
    if (rc) return rc;
    WideBufs &w = g_wide;
    if (hipMemsetAsync(w.misc, 0, 72 * 4, s) != hipSuccess) return GQ_ERR_DEVICE;
    const uint64_t slots = (uint64_t)memo.mask + 1ull;
    hipLaunchKernelGGL(ks_remap, dim3((uint32_t)std::min<uint64_t>((slots + WBLOCK - 1) / WBLOCK, 4096)),
                       dim3(WBLOCK), 0, s, memo, snap_wtab(r.out), pm, w.misc);
    if (hipMemcpyAsync(w.h_misc, w.misc, 72 * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_error("memo snapshot: copying the memo failed");
        return GQ_ERR_DEVICE;
    }
    if (w.h_misc[WERR] || w.h_misc[0] != total) {
        set_error("memo snapshot: %u of %llu sets copied (error bits %u)", w.h_misc[0],
                  (unsigned long long)total, w.h_misc[WERR]);
        return GQ_ERR_DEVICE;
    }
    r.out.n_sets = total;
    r.done = true;
    return GQ_OK;
}

int snapshot_walk(const SnapTable &t, uint64_t root, std::vector<WideNode> &nodes) {
    const int k = __builtin_popcountll(root);
    if (!t.tab || !root || (t.n_rels < 64 && (root >> t.n_rels))) {
        set_error("not a set of the snapshot's %d relations", t.n_rels);
        return GQ_ERR_INVALID;
    }
    return with_device_stream([&](int device, void *sv) -> int {
        if (device != t.device) {
            set_error("the snapshot lives on device %d, the library on %d", t.device, device);
            return GQ_ERR_UNSUPPORTED;
        }
        hipStream_t s = (hipStream_t)sv;
        WideBufs &w = g_wide;
        int rc = wide_init(w, device);
        if (rc) return rc;
        WCHK(hipMemsetAsync(w.misc, 0, 72 * 4, s));
        hipLaunchKernelGGL(kw_walk, dim3(1), dim3(WW), 0, s, snap_wtab(t), (uint64_t)root, w.nodes, w.misc);
        WCHK(hipGetLastError());
        WCHK(hipMemcpyAsync(w.h_misc, w.misc, 72 * 4, hipMemcpyDeviceToHost, s));
        WCHK(hipMemcpyAsync(w.h_nodes, w.nodes, (size_t)(2 * k - 1) * sizeof(WideNode), hipMemcpyDeviceToHost, s));
        WCHK(hipStreamSynchronize(s));
        if (w.h_misc[WERR] || w.h_misc[WERR + 1] != (uint32_t)(2 * k - 1)) {
            set_error("relation set %#llx is not a connected set of the snapshot",
                      (unsigned long long)root);
            return GQ_ERR_INVALID;
        }
        nodes.assign(w.h_nodes, w.h_nodes + (2 * k - 1));
        return GQ_OK;
    });
}

int snapshot_read(const SnapTable &t, std::vector<SnapRecord> &recs) {
    recs.clear();
    if (!t.tab) {
        set_error("empty snapshot");
        return GQ_ERR_INVALID;
    }
    std::vector<WSlot> h((size_t)1 << t.lg);
    int rc = with_device_stream([&](int device, void *sv) -> int {
        if (device != t.device) {
            set_error("the snapshot lives on device %d, the library on %d", t.device, device);
            return GQ_ERR_UNSUPPORTED;
        }
        hipStream_t s = (hipStream_t)sv;
        WCHK(hipMemcpyAsync(h.data(), t.tab, h.size() * sizeof(WSlot), hipMemcpyDeviceToHost, s));
        WCHK(hipStreamSynchronize(s));
        return GQ_OK;
    });
    if (rc) return rc;
    recs.reserve(t.n_sets);
    for (const WSlot &x : h) {
        if (!x.key) continue;
        SnapRecord r;
        r.set = x.key;
        r.left = x.left;
        r.rows = x.rec.x;
        r.startup = x.rec.y;
        r.total = x.rec.z;
        uint32_t wb;
        memcpy(&wb, &x.rec.w, 4);
        r.width = (wb >> 31) ? (int32_t)(wb & 0x1fffffu) : (int32_t)wb;   /* rec_width */
        recs.push_back(r);
    }
    std::sort(recs.begin(), recs.end(), [](const SnapRecord &a, const SnapRecord &b) {
        const int pa = __builtin_popcountll(a.set), pb = __builtin_popcountll(b.set);
        return pa != pb ? pa < pb : a.set < b.set;
    });
    return GQ_OK;
}

int snapshot_build(const std::vector<SnapRecord> &recs, int n_rels, SnapTable &t) {
    std::vector<WSlot> h(recs.size());
    for (size_t i = 0; i < recs.size(); i++) {
        h[i].key = recs[i].set;
        h[i].left = recs[i].left;
        h[i].rec = make_float4(recs[i].rows, recs[i].startup, recs[i].total, 0.0f);
        const uint32_t wb = (uint32_t)recs[i].width;
        memcpy(&h[i].rec.w, &wb, 4);
    }
    return with_device_stream([&](int device, void *sv) -> int {
        hipStream_t s = (hipStream_t)sv;
        WideBufs &w = g_wide;
        int rc = wide_init(w, device);
        if (rc) return rc;
        if ((rc = snap_alloc(t, recs.size(), n_rels, device, s))) return rc;
        WSlot *d = nullptr;
        if (hipMalloc((void **)&d, std::max<size_t>(1, h.size()) * sizeof(WSlot)) != hipSuccess) {
            (void)hipGetLastError();
            snapshot_free(t);
            set_error("hipMalloc for the imported records failed");
            return GQ_ERR_NOMEM;
        }
        bool ok = hipMemcpyAsync(d, h.data(), h.size() * sizeof(WSlot), hipMemcpyHostToDevice, s) == hipSuccess &&
                  hipMemsetAsync(w.misc, 0, 72 * 4, s) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(ks_fill, dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(
                                         (h.size() + WBLOCK - 1) / WBLOCK, 4096))),
                               dim3(WBLOCK), 0, s, snap_wtab(t), (const WSlot *)d, (uint64_t)h.size(), w.misc);
            ok = hipGetLastError() == hipSuccess &&
                 hipMemcpyAsync(w.h_misc, w.misc, 72 * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess;
        }
        (void)hipFree(d);
        if (!ok) {
            snapshot_free(t);
            set_error("memo snapshot: building the table failed");
            return GQ_ERR_DEVICE;
        }
        if (w.h_misc[WERR]) {
            snapshot_free(t);
            set_error("memo snapshot: %s", (w.h_misc[WERR] & WERR_PROBE) ? "a set appears twice"
                                                                          : "the table is full");
            return GQ_ERR_INVALID;
        }
        t.n_sets = h.size();
        return GQ_OK;
    });
}

template <bool DPSUB, bool FAST>
static void launch_level(const WideHdr &h, uint32_t grid, size_t lds, hipStream_t s, const uint32_t *blob, WTab t,
                         uint32_t *lst, int k, uint32_t off_k, uint32_t n_k, uint32_t off_n, uint32_t cap_n,
                         uint32_t *misc) {
    hipLaunchKernelGGL((kw_level<DPSUB, FAST>), dim3(grid), dim3(WBLOCK), lds, s, h, blob, t, lst, k, off_k, n_k,
                       off_n, cap_n, misc);
}

/* A wide DP over a graph with cycles: the level sizes are learnt one level
 * at a time (the host reads each level's count before launching it), in a
 * memo of 2^lg 32-byte slots and a list of 2^(lg-1) entries sized from
 * gpuqo_max_memo_size_mb; a DP that outgrows them fails with GQ_ERR_NOMEM.
 * The interrupt flag is tested before every level. */
static int run_wide_general(const WideInput &in, DPAlgo algo, int max_memo_mb, std::vector<WideNode> &nodes,
                            DPStats &st) {
    const WideHdr &h = in.hdr;
    const int n = h.n_rels;
    const size_t lds = (((size_t)h.words * 4 + 15) & ~(size_t)15) + (WBLOCK / WW) * sizeof(WgScratch);
    if (lds > 64 * 1024) {
        set_error("query description too large for LDS (%zu bytes)", lds);
        return GQ_ERR_UNSUPPORTED;
    }
    const uint64_t budget = (uint64_t)std::max(max_memo_mb, 1) << 20;
    uint32_t lg = 8;
    while (lg < 31 && (32ull << (lg + 1)) + (4ull << lg) <= budget) lg++;
    const uint64_t slots = 1ull << lg, list_cap = slots / 2;
    return with_device_stream([&](int device, void *sv) -> int {
        hipStream_t s = (hipStream_t)sv;
        WideBufs &w = g_wide;
        int rc = wide_init(w, device);
        if (rc) return rc;
        if ((rc = wensure(&w.tab, w.cap_tab, (size_t)slots))) return rc;
        if ((rc = wensure(&w.lst, w.cap_lst, (size_t)list_cap))) return rc;
        if ((rc = wensure(&w.blob, w.cap_blob, in.blob.size()))) return rc;
        const double t_first = host_now_ms();
        WCHK(hipMemcpyAsync(w.blob, in.blob.data(), in.blob.size() * 4, hipMemcpyHostToDevice, s));
        WCHK(hipMemsetAsync(w.tab, 0, sizeof(WSlot) * slots, s));
        WCHK(hipMemsetAsync(w.misc, 0, 72 * 4, s));
        WCHK(hipEventRecord(w.e0, s));
        WTab t;
        t.s = w.tab;
        t.lg = lg;
        t.mask = (uint32_t)(slots - 1u);
        hipLaunchKernelGGL(kw_gbase, dim3(1), dim3(WW), 0, s, h, (const uint32_t *)w.blob, t, w.lst,
                           (uint32_t)(list_cap - (uint64_t)n), w.misc);
        int launches = 1;
        uint64_t off_k = (uint64_t)n, sets = 0;
        bool stopped = false;
        for (int k = 2; k <= n; k++) {
            WCHK(hipMemcpyAsync(w.h_misc, w.misc, 72 * 4, hipMemcpyDeviceToHost, s));
            WCHK(hipStreamSynchronize(s));
            if (w.h_misc[WERR]) {          /* reported below; no further level or walk */
                stopped = true;
                break;
            }
            if (interrupted()) {
                set_error("interrupted before level %d", k);
                return GQ_ERR_INTERRUPTED;
            }
            const uint32_t n_k = w.h_misc[k];
            if (n_k == 0 || (k == n && n_k != 1)) {
                set_error("wide DP: %u connected sets of size %d", n_k, k);
                return GQ_ERR_DEVICE;
            }
            sets += n_k;
            const uint64_t off_n = off_k + n_k;
            const uint32_t cap_n = k < n ? (uint32_t)(list_cap - std::min<uint64_t>(off_n, list_cap)) : 0u;
            const uint32_t grid = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(((uint64_t)n_k + WBLOCK / WW - 1) / (WBLOCK / WW), WGRID_MAX));
            if (algo == DP_DPSUB)
                hipLaunchKernelGGL(kw_glevel<true>, dim3(grid), dim3(WBLOCK), lds, s, h, (const uint32_t *)w.blob, t,
                                   w.lst, k, (uint32_t)off_k, n_k, (uint32_t)off_n, cap_n, w.misc);
            else
                hipLaunchKernelGGL(kw_glevel<false>, dim3(grid), dim3(WBLOCK), lds, s, h, (const uint32_t *)w.blob,
                                   t, w.lst, k, (uint32_t)off_k, n_k, (uint32_t)off_n, cap_n, w.misc);
            launches++;
            off_k = off_n;
        }
        if (!stopped) hipLaunchKernelGGL(kw_walk, dim3(1), dim3(WW), 0, s, t, full_set(n), w.nodes, w.misc);
        WCHK(hipGetLastError());
        WCHK(hipEventRecord(w.e1, s));
        const double t_last = host_now_ms();
        WCHK(hipMemcpyAsync(w.h_misc, w.misc, 72 * 4, hipMemcpyDeviceToHost, s));
        WCHK(hipMemcpyAsync(w.h_nodes, w.nodes, (size_t)(2 * n - 1) * sizeof(WideNode), hipMemcpyDeviceToHost, s));
        WCHK(hipStreamSynchronize(s));
        const double t_synced = host_now_ms();
        const uint32_t err = w.h_misc[WERR];
        if (err & WERR_FULL) {
            set_error("the DP's connected sets outgrow the memo of gpuqo_max_memo_size_mb (%d)", max_memo_mb);
            return GQ_ERR_NOMEM;
        }
        if (err & WERR_BIG) {
            set_error("a set of the DP has a biconnected block of more than %d relations that is not a simple "
                      "cycle (or, under MPDP, a cycle of all 64); plan it with gpuqo_idp_type / "
                      "gpuqo_idp_n_iters", WG_BLOCK_MAX);
            return GQ_ERR_UNSUPPORTED;
        }
        if (err) {
            set_error("wide DP memo inconsistency (error bits %u)", err);
            return GQ_ERR_DEVICE;
        }
        if (w.h_misc[WERR + 1] != (uint32_t)(2 * n - 1)) {
            set_error("wide DP plan walk returned %u nodes", w.h_misc[WERR + 1]);
            return GQ_ERR_DEVICE;
        }
        nodes.assign(w.h_nodes, w.h_nodes + (2 * n - 1));
        float ms = 0.0f;
        WCHK(hipEventElapsedTime(&ms, w.e0, w.e1));
        st = DPStats();
        st.candidates = sets;
        st.sets = sets;
        st.pairs = (uint64_t)w.h_misc[WPAIRS] | ((uint64_t)w.h_misc[WPAIRS + 1] << 32);
        st.memo_bytes = (32ull << lg) + 4ull * list_cap;
        st.memo_kind = GQ_MEMO_HASH;
        st.gpu_ms = ms;
        st.eval_ms = ms;
        st.eval_launches = launches;
        st.t_first_launch = t_first;
        st.t_last_launch = t_last;
        st.t_synced = t_synced;
        if (g_snap_req && g_snap_req->n == n) return wide_snapshot(t, sets + (uint64_t)n, n, device, s);
        return GQ_OK;
    });
}


int run_wide_dp(const WideInput &in, DPAlgo algo, int max_memo_mb, std::vector<WideNode> &nodes, DPStats &st) {
    const WideHdr &h = in.hdr;
    const int n = h.n_rels;
    if (algo == DP_DPSIZE) {
        set_error("DPsize runs on the GPU for at most %d relations (got %d)", GQ_MAX_DP_RELS, n);
        return GQ_ERR_UNSUPPORTED;
    }
    if (!h.is_tree) return run_wide_general(in, algo, max_memo_mb, nodes, st);
    const size_t lds = ((size_t)h.words * 4 + 15) & ~(size_t)15;
    if (lds > 64 * 1024) {
        set_error("query description too large for LDS (%zu bytes)", lds);
        return GQ_ERR_UNSUPPORTED;
    }
    uint64_t total = 0, pairs = 0;
    for (int k = 1; k <= n; k++) {
        total += in.counts[k];
        pairs += 2ull * (uint64_t)(k - 1) * in.counts[k];
    }
    const uint64_t bytes = wide_dp_bytes(in);
    if (bytes > ((uint64_t)std::max(max_memo_mb, 1) << 20) || total >= (1ull << 30)) {
        set_error("the DP's %llu connected sets need %llu MB of memo, above gpuqo_max_memo_size_mb (%d)",
                  (unsigned long long)total, (unsigned long long)(bytes >> 20), max_memo_mb);
        return GQ_ERR_NOMEM;
    }
    const uint32_t lg = wide_slots_log2(in);
    std::vector<uint32_t> off(n + 2, 0);                  /* level k's list starts at off[k] */
    for (int k = 2; k <= n + 1; k++) off[k] = off[k - 1] + (uint32_t)in.counts[k - 1];
    return with_device_stream([&](int device, void *sv) -> int {
        hipStream_t s = (hipStream_t)sv;
        WideBufs &w = g_wide;
        int rc = wide_init(w, device);
        if (rc) return rc;
        if ((rc = wensure(&w.tab, w.cap_tab, (size_t)1 << lg))) return rc;
        if ((rc = wensure(&w.lst, w.cap_lst, (size_t)total))) return rc;
        if ((rc = wensure(&w.blob, w.cap_blob, in.blob.size()))) return rc;
        const double t_first = host_now_ms();
        WCHK(hipMemcpyAsync(w.blob, in.blob.data(), in.blob.size() * 4, hipMemcpyHostToDevice, s));
        WCHK(hipMemsetAsync(w.tab, 0, sizeof(WSlot) << lg, s));
        WCHK(hipMemsetAsync(w.misc, 0, 72 * 4, s));
        WCHK(hipEventRecord(w.e0, s));
        WTab t;
        t.s = w.tab;
        t.lg = lg;
        t.mask = (uint32_t)(((uint64_t)1 << lg) - 1u);
        hipLaunchKernelGGL(kw_base, dim3(1), dim3(WW), 0, s, h, (const uint32_t *)w.blob, t, w.lst, off[2],
                           (uint32_t)in.counts[2], w.misc);
        int launches = 1;
        for (int k = 2; k <= n; k++) {
            const uint32_t n_k = (uint32_t)in.counts[k];
            const uint32_t spw = WW / (uint32_t)(k - 1);
            const uint64_t waves = (n_k + spw - 1) / spw;
            const uint32_t grid = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>((waves + WBLOCK / WW - 1) / (WBLOCK / WW), WGRID_MAX));
            const uint32_t off_n = k < n ? off[k + 1] : 0u, cap_n = k < n ? (uint32_t)in.counts[k + 1] : 0u;
            const bool dpsub = algo == DP_DPSUB, fast = h.tree_fast != 0;
            if (dpsub && fast)
                launch_level<true, true>(h, grid, lds, s, w.blob, t, w.lst, k, off[k], n_k, off_n, cap_n, w.misc);
            else if (dpsub)
                launch_level<true, false>(h, grid, lds, s, w.blob, t, w.lst, k, off[k], n_k, off_n, cap_n, w.misc);
            else if (fast)
                launch_level<false, true>(h, grid, lds, s, w.blob, t, w.lst, k, off[k], n_k, off_n, cap_n, w.misc);
            else
                launch_level<false, false>(h, grid, lds, s, w.blob, t, w.lst, k, off[k], n_k, off_n, cap_n, w.misc);
            launches++;
        }
        hipLaunchKernelGGL(kw_walk, dim3(1), dim3(WW), 0, s, t, full_set(n), w.nodes, w.misc);
        WCHK(hipGetLastError());
        WCHK(hipEventRecord(w.e1, s));
        const double t_last = host_now_ms();
        WCHK(hipMemcpyAsync(w.h_misc, w.misc, 72 * 4, hipMemcpyDeviceToHost, s));
        WCHK(hipMemcpyAsync(w.h_nodes, w.nodes, (size_t)(2 * n - 1) * sizeof(WideNode), hipMemcpyDeviceToHost, s));
        WCHK(hipStreamSynchronize(s));
        const double t_synced = host_now_ms();
        const uint32_t err = w.h_misc[WERR];
        if (err) {
            set_error("wide DP memo inconsistency (error bits %u)", err);
            return GQ_ERR_DEVICE;
        }
        for (int k = 3; k <= n; k++)
            if (w.h_misc[k] != (uint32_t)in.counts[k]) {
                set_error("wide DP generated %u sets of size %d, expected %llu", w.h_misc[k], k,
                          (unsigned long long)in.counts[k]);
                return GQ_ERR_DEVICE;
            }
        if (w.h_misc[2] != (uint32_t)in.counts[2] || w.h_misc[WERR + 1] != (uint32_t)(2 * n - 1)) {
            set_error("wide DP plan walk returned %u nodes", w.h_misc[WERR + 1]);
            return GQ_ERR_DEVICE;
        }
        nodes.assign(w.h_nodes, w.h_nodes + (2 * n - 1));
        float ms = 0.0f;
        WCHK(hipEventElapsedTime(&ms, w.e0, w.e1));
        st = DPStats();
        st.candidates = total - in.counts[1];
        st.sets = total - in.counts[1];
        st.pairs = pairs;
        st.memo_bytes = bytes;
        st.memo_kind = GQ_MEMO_HASH;
        st.gpu_ms = ms;
        st.eval_ms = ms;
        st.eval_launches = launches;
        st.t_first_launch = t_first;
        st.t_last_launch = t_last;
        st.t_synced = t_synced;
        if (g_snap_req && g_snap_req->n == n) return wide_snapshot(t, total, n, device, s);
        return GQ_OK;
    });
}

}  // namespace gq
