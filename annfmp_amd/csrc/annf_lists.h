// Sorted key lists kept in registers by the full-patch leaf kernels and the
// exact field: every index is static, so a list lives in VGPRs.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace annf {

__device__ __forceinline__ uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm volatile("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Sorted list of LM 32-bit keys (ascending; a list whose live part is its
// top K slots holds 0 in the front slots, which never move).  Slot s of the
// list with `key` added is the median of (old l[s-1], key, old l[s]); a key
// above every entry changes nothing.  Lists longer than 8 run the network
// only when the key beats the last entry.
template <int LM>
__device__ __forceinline__ void list_insert(uint32_t (&l)[LM], uint32_t key) {
    if constexpr (LM <= 8) {
#pragma unroll
        for (int s = LM - 1; s >= 1; --s) l[s] = med3_u32(l[s - 1], key, l[s]);
        l[0] = min(l[0], key);
    } else if (key < l[LM - 1]) {
#pragma unroll
        for (int s = LM - 1; s >= 1; --s) l[s] = med3_u32(l[s - 1], key, l[s]);
        l[0] = min(l[0], key);
    }
}

// The same for 64-bit keys: l[s] <= key holds for a prefix of the slots;
// slot s keeps its key inside that prefix, takes `key` right after it and
// its left neighbour's key further on.
template <int LM>
__device__ __forceinline__ void list_insert64(uint64_t (&l)[LM], uint64_t key) {
    bool le[LM];
#pragma unroll
    for (int s = 0; s < LM; ++s) le[s] = l[s] <= key;
#pragma unroll
    for (int s = LM - 1; s >= 1; --s) l[s] = le[s] ? l[s] : (le[s - 1] ? key : l[s - 1]);
    l[0] = le[0] ? l[0] : key;
}

}  // namespace annf
