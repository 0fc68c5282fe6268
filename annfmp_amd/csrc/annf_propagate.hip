// The propagated pass of the full-patch leaf k-NN (the parallel form of
// base.c:255-282) and its helpers: the (query, leaf) pairs that the first
// pass's lists of the row above propagate, each leaf's points in ascending B
// index (the position order of the leaf passes), and the merge of a query's
// first-pass and propagated keys into the field.
//
// Keys are (squared distance << 32) | B index, so ties go to the lower B
// index, the order of the exact field.  The first list entry is the match.
#include "annf_internal.h"

namespace annf {

namespace {

constexpr uint64_t kNoKey = ~0ull;

// Sorted key list of LM slots whose live part is the top K slots; the front
// slots hold 0, which never moves (an insert lands after every key <= it),
// so every index is static.
// Branch-free: the list is sorted, so l[s] <= key holds for a prefix of the
// slots; slot s keeps its key inside that prefix, takes `key` right after it
// and its left neighbour's key further on (a key above every entry changes
// nothing).
template <int LM>
__device__ __forceinline__ void klist_insert(uint64_t (&l)[LM], uint64_t key) {
    bool le[LM];
#pragma unroll
    for (int s = 0; s < LM; ++s) le[s] = l[s] <= key;
#pragma unroll
    for (int s = LM - 1; s >= 1; --s) l[s] = le[s] ? l[s] : (le[s - 1] ? key : l[s - 1]);
    l[0] = le[0] ? l[0] : key;
}

__device__ __forceinline__ float sqrt_f(uint32_t d2) {
    // correctly rounded: the FP64 root rounded once more to FP32 (2*24+2 <= 53)
    return (float)__builtin_sqrt((double)d2);
}

// Propagated leaves (the parallel form of base.c:255-282): for A patch q
// below the first row and k < prop_k, the B patch one row below the k-th
// match of the patch above (first-pass keys), its leaf pli[.]; leaves the
// first pass already scanned (natleaf[q] ^ r, r < nprobe) or equal to an
// earlier k's are dropped (-1).  Counts the kept (q, k) pairs per leaf.
__global__ void k_prop_pairs(const uint64_t *__restrict__ keys1, int q0, int nq, int K, int prop_k, int nprobe,
                             const int *__restrict__ natleaf, const int *__restrict__ pli, int n_cols, int n_b,
                             int *__restrict__ pairleaf, int *__restrict__ count) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nq * prop_k) return;
    const int ql = (int)(t / prop_k), k = (int)(t - (long long)ql * prop_k), q = q0 + ql;
    int leaf = -1;
    if (q >= n_cols) {
        auto leaf_of = [&](int kk) -> int {
            const uint64_t key = keys1[(size_t)(q - n_cols) * K + kk];
            if (key == kNoKey) return -1;
            int o = (int)(uint32_t)key;
            if (o + n_cols < n_b) o += n_cols;
            return pli[o];
        };
        leaf = leaf_of(k);
        if (leaf >= 0 && (leaf ^ natleaf[q]) < nprobe) leaf = -1;
        for (int kk = 0; kk < k && leaf >= 0; ++kk)
            if (leaf_of(kk) == leaf) leaf = -1;
    }
    pairleaf[t] = leaf;
    if (count && leaf >= 0) wave_count1(count, leaf);
}

// Each leaf's points in ascending B index (the BF16 kernel's tile position
// order), by rank: out[from + #(smaller in the leaf)] = index.
constexpr int kLeafSortLds = 2048;
constexpr int kLeafSortTPB = 128;   // a leaf of the usual size: one point per thread
__global__ __launch_bounds__(kLeafSortTPB) void k_leaf_order(const int *__restrict__ leaves, const int *__restrict__ orig,
                                                    int *__restrict__ out) {
    __shared__ int v[kLeafSortLds];
    const int l = blockIdx.x, from = leaves[2 * l], c = leaves[2 * l + 1] - from;
    const int *src = orig + from;
    if (c <= kLeafSortLds) {
        for (int i = threadIdx.x; i < c; i += kLeafSortTPB) v[i] = src[i];
        __syncthreads();
        for (int i = threadIdx.x; i < c; i += kLeafSortTPB) {
            const int x = v[i];
            int r = 0;
            for (int j = 0; j < c; ++j) r += v[j] < x;
            out[from + r] = x;
        }
    } else {
        for (int i = threadIdx.x; i < c; i += kLeafSortTPB) {
            const int x = src[i];
            int r = 0;
            for (int j = 0; j < c; ++j) r += src[j] < x;
            out[from + r] = x;
        }
    }
}

// The K smallest of a query's first-pass keys and its propagated pairs'
// keys (k_prop_pairs leaves out every leaf the first pass scanned and
// repeats, so the leaves are disjoint and no B index comes twice) -> knn,
// nn, dist.  One thread per query; a wavefront's 64 queries own one
// contiguous run of pair keys, which the wavefront reads 16 keys per query
// at a time, coalesced, through LDS (each thread then takes its own row).
constexpr int kFinChunk = 16;

// orders the LDS accesses of the lanes of one wavefront
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int KM>
__global__ __launch_bounds__(256) void k_finish(const uint64_t *__restrict__ keys1,
                                                const uint64_t *__restrict__ keys2,
                                                const int *__restrict__ pairleaf, int q0, int nq, int K,
                                                int prop_k, int *__restrict__ knn, int *__restrict__ nn,
                                                float *__restrict__ dist, uint64_t *__restrict__ keys_out) {
    __shared__ uint64_t tile[4][kWave][kFinChunk + 1];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int wq0 = blockIdx.x * blockDim.x + w * kWave;    // the wavefront's first query (local)
    if (wq0 >= nq) return;                                   // whole wavefronts leave together
    const int ql = wq0 + lane;
    const bool live = ql < nq;
    const int q = q0 + (live ? ql : wq0);
    // the first-pass keys are ascending already: they are the list
    uint64_t l[KM];
#pragma unroll
    for (int s = 0; s < KM; ++s) l[s] = s < KM - K ? 0ull : keys1[(size_t)q * K + (s - (KM - K))];
    if (keys2) {
        // pairs of leaf -1 were not scanned: their keys are skipped
        uint32_t dead = 0;
        for (int k = 0; k < prop_k; ++k)
            if (live && pairleaf[(size_t)ql * prop_k + k] < 0) dead |= 1u << k;
        const int n2 = prop_k * K;
        const int nw = min(kWave, nq - wq0);                 // queries of this wavefront
        const uint64_t *src = keys2 + (size_t)wq0 * n2;
        for (int c0 = 0; c0 < n2; c0 += kFinChunk) {
            // keys c0 .. c0 + 15 of every query of the wavefront: eight lanes
            // per query row, two keys each, eight rows per pass
            for (int e = lane; e < nw * (kFinChunk / 2); e += kWave) {
                const int row = e / (kFinChunk / 2), u = 2 * (e - row * (kFinChunk / 2));
                const size_t o = (size_t)row * n2 + c0 + u;
                tile[w][row][u] = c0 + u < n2 ? src[o] : kNoKey;
                tile[w][row][u + 1] = c0 + u + 1 < n2 ? src[o + 1] : kNoKey;
            }
            wave_lds_sync();
#pragma unroll
            for (int u = 0; u < kFinChunk; ++u) {
                const int s = c0 + u;
                const uint64_t key = (s < n2 && !((dead >> (s / K)) & 1u)) ? tile[w][lane][u] : kNoKey;
                if (key < l[KM - 1]) klist_insert(l, key);
            }
            wave_lds_sync();
        }
    }
    if (!live) return;
#pragma unroll
    for (int s = 0; s < KM; ++s) {
        if (s < KM - K) continue;
        const uint64_t key = l[s];
        const int j = key == kNoKey ? -1 : (int)(uint32_t)key;
        const int r = s - (KM - K);
        if (knn) knn[(size_t)q * K + r] = j;
        if (keys_out) keys_out[(size_t)q * K + r] = key;
        if (r == 0) {
            if (nn) nn[q] = j;
            if (dist) dist[q] = key == kNoKey ? kFltMax : sqrt_f((uint32_t)(key >> 32));
        }
    }
}

}  // namespace

hipError_t prop_pairs(const uint64_t *keys1, int q0, int nq, int K, int prop_k, int nprobe, const int *natleaf,
                      const int *pli, int n_cols, int n_b, int *pairleaf, int *count, hipStream_t st) {
    const long long n = (long long)nq * prop_k;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prop_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys1, q0, nq, K, prop_k,
                       nprobe, natleaf, pli, n_cols, n_b, pairleaf, count);
    return hipGetLastError();
}

hipError_t finish_field(const uint64_t *keys1, const uint64_t *keys2, const int *pairleaf, int q0, int nq, int K,
                        int prop_k, int *knn, int *nn, float *dist, uint64_t *keys_out, hipStream_t st) {
    if (keys2 && !pairleaf) return hipErrorInvalidValue;
    if (nq == 0) return hipSuccess;
    const dim3 g((unsigned)((nq + 255) / 256)), b(256);
    if (K <= 8)
        hipLaunchKernelGGL(k_finish<8>, g, b, 0, st, keys1, keys2, pairleaf, q0, nq, K, prop_k, knn, nn, dist,
                           keys_out);
    else
        hipLaunchKernelGGL(k_finish<32>, g, b, 0, st, keys1, keys2, pairleaf, q0, nq, K, prop_k, knn, nn, dist,
                           keys_out);
    return hipGetLastError();
}

hipError_t leaf_point_order(const TreeView &t, int *out, hipStream_t st) {
    hipLaunchKernelGGL(k_leaf_order, dim3(1u << t.depth), dim3(kLeafSortTPB), 0, st, t.leaves, t.orig, out);
    return hipGetLastError();
}

}  // namespace annf
