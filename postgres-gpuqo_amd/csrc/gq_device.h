/*
 * gq_device.h -- device helpers shared by the DP kernels of gq_dp.hip
 * (32-bit relation sets) and gq_wide.hip (64-bit relation sets): the memo
 * record format, the order-preserving cost key and the open-addressing
 * table of 64-bit sets.
 */
#ifndef GQ_DEVICE_H
#define GQ_DEVICE_H

#include <hip/hip_runtime.h>

#include "gq_cost_model.h"

namespace gq {

/* Memo record (float4): rows, startup, total and a width word, which is the
 * width itself or, bit 31 set, hash_pack() of the relation (the width in
 * its low 21 bits).  Widths are below 2^31, so bit 31 tells the two forms
 * apart. */
__device__ __forceinline__ int rec_width(uint32_t w) { return (w >> 31) ? (int)(w & 0x1fffffu) : (int)w; }

__device__ __forceinline__ Rel rel_of(float4 v) {
    Rel r;
    r.rows = v.x;
    r.startup = v.y;
    r.total = v.z;
    const uint32_t w = __float_as_uint(v.w);
    r.width = rec_width(w);
    r.hs = (w >> 31) ? w : 0u;
    return r;
}

/* pack = false keeps the plain width (readers then size the hash table per
 * join). */
__device__ __forceinline__ float4 record_of(const Rel &r, int work_mem, bool pack = true) {
    const uint32_t hs = pack ? hash_pack(r.rows, r.width, work_mem) : 0u;
    return make_float4(r.rows, r.startup, r.total, __uint_as_float(hs ? hs : (uint32_t)r.width));
}

/* unsigned order of non-NaN floats; -0 and +0 map to the same key */
__device__ __forceinline__ uint32_t ord_key(float x) {
    if (x == 0.0f) x = 0.0f;
    uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* 32-byte slot of an open-addressing memo over 64-bit relation sets (the
 * wide DP's memo, and memo snapshots): key 0 is empty; left is the outer
 * side of the chosen split (0 for base relations); rec as above. */
struct alignas(16) WSlot {
    unsigned long long key;    /* the set; 0 = empty */
    unsigned long long left;   /* outer side of the chosen split (0 for base relations) */
    float4 rec;                /* rows, startup, total, width word (rel_of / record_of) */
};
static_assert(sizeof(WSlot) == 32, "32-byte memo slots");

struct WTab {
    WSlot *s;
    uint32_t mask, lg;
};

__device__ __forceinline__ uint32_t wslot0(uint64_t key, uint32_t lg) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64u - lg));
}

/* slot of a key inserted by an earlier launch; ~0u if it is missing */
__device__ __forceinline__ uint32_t wfind(const WTab &t, uint64_t key) {
    uint32_t h = wslot0(key, t.lg);
    for (uint32_t i = 0; i <= t.mask; i++) {
        const unsigned long long k = t.s[h].key;
        if (k == key) return h;
        if (k == 0ull) break;
        h = (h + 1u) & t.mask;
    }
    return ~0u;
}

/* claims a slot for key (each set is generated once); ~0u if the table is full */
__device__ __forceinline__ uint32_t winsert(const WTab &t, uint64_t key) {
    uint32_t h = wslot0(key, t.lg);
    for (uint32_t i = 0; i <= t.mask; i++) {
        const unsigned long long old = atomicCAS(&t.s[h].key, 0ull, (unsigned long long)key);
        if (old == 0ull || old == key) return h;
        h = (h + 1u) & t.mask;
    }
    return ~0u;
}

/* claims the slot of key; *fresh when this call inserted it */
__device__ __forceinline__ uint32_t winsert_fresh(const WTab &t, uint64_t key, bool *fresh) {
    uint32_t h = wslot0(key, t.lg);
    for (uint32_t i = 0; i <= t.mask; i++) {
        const unsigned long long old = atomicCAS(&t.s[h].key, 0ull, (unsigned long long)key);
        if (old == 0ull || old == key) {
            *fresh = old == 0ull;
            return h;
        }
        h = (h + 1u) & t.mask;
    }
    *fresh = false;
    return ~0u;
}

}  // namespace gq

#endif
