// Bucketing by leaf as a stable radix sort of (leaf, index) pairs (rocPRIM's
// device radix sort): the order within a leaf is the index order, and the
// per-leaf offsets are read off the sorted keys -- no counting atomics.
#include <cstring>

#include "annf_internal.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace annf {

namespace {

// keys for the sort: leaf, or nleaves for an entry without one (leaf < 0)
__global__ void k_sort_keys(const int *__restrict__ leaf, int n, int nleaves, unsigned *__restrict__ key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int l = leaf[i];
        key[i] = (unsigned)(l < 0 ? nleaves : l);
    }
}

// offset[l] = first sorted position with key >= l, for l = 0..nleaves
__global__ void k_offsets_from_sorted(const unsigned *__restrict__ key, int n, int nleaves, int *__restrict__ offset) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0) {
        if (i <= nleaves) offset[i] = 0;
        return;
    }
    if (i >= n) return;
    const int k = (int)min(key[i], (unsigned)nleaves);
    const int prev = i > 0 ? (int)min(key[i - 1], (unsigned)nleaves) : -1;
    for (int l = prev + 1; l <= k; ++l) offset[l] = i;
    if (i == n - 1)
        for (int l = k + 1; l <= nleaves; ++l) offset[l] = n;
}

int key_bits(int nleaves) {
    int b = 1;
    while ((1 << b) <= nleaves) ++b;   // keys 0..nleaves inclusive
    return b;
}

}  // namespace

size_t sort_scratch_bytes(int n, int nleaves) {
    size_t bytes = 0;
    rocprim::radix_sort_pairs(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr, (const int *)nullptr,
                              (int *)nullptr, (unsigned int)n, 0u, (unsigned int)key_bits(nleaves));
    return bytes;
}

hipError_t bucket_sorted(const int *leaf, int n, int nleaves, int chunk, int *keys_a_, int *keys_b_, int *order,
                         int *offset, int *chunk_cnt, int *chunk_off, void *scratch, size_t scratch_bytes,
                         hipStream_t st) {
    unsigned *keys_a = reinterpret_cast<unsigned *>(keys_a_);
    unsigned *keys_b = reinterpret_cast<unsigned *>(keys_b_);
    if (n > 0) {
        if (sort_scratch_bytes(n, nleaves) > scratch_bytes) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_sort_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, leaf, n, nleaves, keys_a);
        hipError_t e = rocprim::radix_sort_pairs(scratch, scratch_bytes, keys_a, keys_b,
                                                 rocprim::counting_iterator<int>(0), order, (unsigned int)n, 0u,
                                                 (unsigned int)key_bits(nleaves), st);
        if (e != hipSuccess) return e;
    }
    const int m = n > 0 ? n : nleaves + 1;
    hipLaunchKernelGGL(k_offsets_from_sorted, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, keys_b, n, nleaves,
                       offset);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return chunk_offsets(offset, nleaves, chunk, chunk_cnt, chunk_off, st);
}

}  // namespace annf
