// Bucketing by leaf as a stable radix sort of (leaf, index) pairs (rocPRIM's
// device radix sort): the order within a leaf is the index order, and the
// per-leaf offsets are read off the sorted keys -- no counting atomics.
#include <algorithm>
#include <cstring>

#include "annf_internal.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace annf {

namespace {

// the sort key of an entry: its leaf, or nleaves without one (leaf < 0)
struct LeafKey {
    unsigned nleaves;
    __host__ __device__ unsigned operator()(int l) const { return l < 0 ? nleaves : (unsigned)l; }
};

// chunks of `chunk` entries in leaf l (0 for l = nleaves, the scan's total)
struct LeafChunks {
    const int *offset;
    int nleaves, chunk;
    __host__ __device__ int operator()(int l) const {
        return l < nleaves ? (offset[l + 1] - offset[l] + chunk - 1) / chunk : 0;
    }
};

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
    size_t sort_bytes = 0, scan_bytes = 0;
    rocprim::radix_sort_pairs(nullptr, sort_bytes,
                              rocprim::make_transform_iterator((const int *)nullptr, LeafKey{(unsigned)nleaves}),
                              (unsigned *)nullptr, rocprim::counting_iterator<int>(0), (int *)nullptr,
                              (unsigned int)n, 0u, (unsigned int)key_bits(nleaves));
    rocprim::exclusive_scan(nullptr, scan_bytes,
                            rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0),
                                                             LeafChunks{nullptr, nleaves, 1}),
                            (int *)nullptr, 0, (size_t)nleaves + 1, rocprim::plus<int>());
    return std::max(sort_bytes, scan_bytes);
}

hipError_t bucket_sorted(const int *leaf, int n, int nleaves, int chunk, int *keys, int *order, int *offset,
                         int *chunk_off, void *scratch, size_t scratch_bytes, hipStream_t st) {
    unsigned *keys_b = reinterpret_cast<unsigned *>(keys);
    if (n > 0) {
        if (sort_scratch_bytes(n, nleaves) > scratch_bytes) return hipErrorInvalidValue;
        hipError_t e = rocprim::radix_sort_pairs(
            scratch, scratch_bytes, rocprim::make_transform_iterator(leaf, LeafKey{(unsigned)nleaves}), keys_b,
            rocprim::counting_iterator<int>(0), order, (unsigned int)n, 0u, (unsigned int)key_bits(nleaves), st);
        if (e != hipSuccess) return e;
    }
    const int m = n > 0 ? n : nleaves + 1;
    hipLaunchKernelGGL(k_offsets_from_sorted, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, keys_b, n, nleaves,
                       offset);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // chunk offsets: exclusive sum of the chunks per leaf, nleaves + 1 entries
    size_t bytes = scratch_bytes;
    return rocprim::exclusive_scan(scratch, bytes,
                                   rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0),
                                                                    LeafChunks{offset, nleaves, chunk}),
                                   chunk_off, 0, (size_t)nleaves + 1, rocprim::plus<int>(), st);
}

}  // namespace annf
