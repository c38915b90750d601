/*
 * gq_device.h -- device helpers shared by the DP kernels of gq_dp.hip
 * (32-bit relation sets) and gq_wide.hip (64-bit relation sets): the memo
 * record format and the order-preserving cost key.
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

}  // namespace gq

#endif
