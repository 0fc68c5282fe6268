// The openclopt backend's semantics on the GPU (annfmp/openclopt/src/futhark/
// *.fut as driven by openclopt/src/base.c:257-520), for gfx950: a padded
// power-of-two tree whose split dimensions come from the ancestors' medians,
// mean-of-two medians, left on <= descent, a stackless exact search for row 0
// (one query per lane: kd-traverse.fut's traverseOnce keeps the path in the
// bits of one int), the row chain over de-duplicated propagated leaves, and
// the best of the eight candidates by exact full-patch distance.
//
// Every FP32 operation is rounded on its own (-ffp-contract=off), distances
// are the sequential sum of (q_j - p_j)^2 (util.fut:19-21), and the sort is a
// stable radix sort on radix_sort_float's key (diku-dk/sorts 0.3.14:
// negative floats with every bit flipped, others with the sign bit set).
#include "annf_internal.h"

#include <cfloat>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

namespace annf {

namespace {

constexpr int kOK = 8;   // driverKNN.fut:16, kk = 8
constexpr int kOclMaxD = 64;

__device__ __forceinline__ uint32_t ocl_fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ocl_unkey(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// point idx of the padded set: rows past m are +inf (buildKDtree.fut:88-89)
__device__ __forceinline__ float ocl_val(const float *X, int m, int d, int idx, int j) {
    return idx < m ? X[(size_t)idx * d + j] : INFINITY;
}

// util.fut:19-21 sumSqrsSeq
__device__ __forceinline__ float ocl_sqdist(const float *__restrict__ q, const float *__restrict__ p, int d) {
    float r = 0.0f;
    for (int j = 0; j < d; ++j) {
        const float z = q[j] - p[j];
        const float zz = z * z;
        r = r + zz;
    }
    return r;
}

// per-dimension bounds of the m real points as ordered keys (lo, hi)
__global__ void k_ocl_bounds(const float *__restrict__ X, int m, int d, uint32_t *lub) {
    __shared__ uint32_t smin[kOclMaxD], smax[kOclMaxD];
    if ((int)threadIdx.x < d) { smin[threadIdx.x] = 0xFFFFFFFFu; smax[threadIdx.x] = 0u; }
    __syncthreads();
    const long long tot = (long long)m * d;
    uint32_t mn = 0xFFFFFFFFu, mx = 0u;
    // a thread's elements share one dimension: the stride is a multiple of d
    const long long stride = (long long)gridDim.x * blockDim.x / d * d;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = (int)(t0 % d);
    if (t0 < (long long)gridDim.x * blockDim.x / d * d)
        for (long long e = t0; e < tot; e += stride) {
            const uint32_t k = ocl_fkey(X[e]);
            mn = min(mn, k);
            mx = max(mx, k);
        }
    atomicMin(&smin[j], mn);
    atomicMax(&smax[j], mx);
    __syncthreads();
    if ((int)threadIdx.x < d) {
        atomicMin(&lub[threadIdx.x], smin[threadIdx.x]);
        atomicMax(&lub[d + threadIdx.x], smax[threadIdx.x]);
    }
}

__global__ void k_ocl_init(uint32_t *lub, int d, int *indir, int mp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d) { lub[i] = 0xFFFFFFFFu; lub[d + i] = 0u; }
    if (i < mp) indir[i] = i;
}

// buildKDtree.fut:32-55 and 69-76 for the nodes of level lev, one thread each:
// the box the ancestors' medians cut from the global bounds (an infinite
// median leaves the bound), the first dimension of largest |ub - lb|, the
// closest ancestor splitting the same dimension.
__global__ void k_ocl_dims(const uint32_t *lub, int d, int lev, int *dims, const float *vals, int *clanc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1 << lev)) return;
    const int node = i + (1 << lev) - 1;
    float lo[kOclMaxD], hi[kOclMaxD];
    for (int j = 0; j < d; ++j) { lo[j] = ocl_unkey(lub[j]); hi[j] = ocl_unkey(lub[d + j]); }
    int anc = 0;
    for (int t = 0; t < lev; ++t) {
        const int k = lev - t - 1;
        const int child = (node + 1 - (1 << k)) >> k;
        const int adim = dims[anc];
        const float med = vals[anc];
        if (!isinf(med)) {
            if ((child & 1) == 0) lo[adim] = med;   // a right child: lower bound
            else hi[adim] = med;
        }
        anc = child;
    }
    int best = -1;
    float bv = -FLT_MAX;
    for (int j = 0; j < d; ++j) {
        const float w = fabsf(hi[j] - lo[j]);
        if (!(bv >= w)) { best = j; bv = w; }
    }
    int c = node, res = -1;
    while (c != 0 && res == -1) {
        const int par = (c - 1) >> 1;
        res = dims[par] == best ? par : -1;
        c = par;
    }
    dims[node] = best;
    clanc[node] = res;
}

// sort keys of level lev: (node << 32) | key of the node's dimension
__global__ void k_ocl_keys(const float *__restrict__ X, int m, int d, int mp, int lev, const int *dims,
                           const int *__restrict__ indir, unsigned long long *keys) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= mp) return;
    const int ppn = mp >> lev, i = p / ppn;
    const int dim = dims[i + (1 << lev) - 1];
    keys[p] = ((unsigned long long)i << 32) | ocl_fkey(ocl_val(X, m, d, indir[p], dim));
}

// buildKDtree.fut:132-135: the mean of the sorted chunk's middle pair
__global__ void k_ocl_medians(const float *__restrict__ X, int m, int d, int mp, int lev, const int *dims,
                              const int *__restrict__ indir, float *vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1 << lev)) return;
    const int node = i + (1 << lev) - 1, ppn = mp >> lev, mi = ppn / 2;
    const int dim = dims[node];
    const float a = ocl_val(X, m, d, indir[i * ppn + mi], dim);
    const float b = ocl_val(X, m, d, indir[i * ppn + mi - 1], dim);
    vals[node] = (a + b) / 2.0f;
}

// leaf rows in tree order (+inf pads) and point -> leaf (buildKDtree.fut:143,
// driverKNN.fut buildKDtree)
__global__ void k_ocl_leaves(const float *__restrict__ X, int m, int d, int mp, int ppl, const int *__restrict__ indir,
                             float *leaf_pts, int *orig2leaf) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)mp * d) return;
    const int p = (int)(e / d), j = (int)(e - (long long)p * d);
    const int idx = indir[p];
    leaf_pts[e] = ocl_val(X, m, d, idx, j);
    if (j == 0) orig2leaf[idx] = p / ppl;
}

__device__ __forceinline__ int ocl_find_leaf(const int *dims, const float *vals, int q, const float *qv) {
    int node = 0;
    while (node < q) node = qv[dims[node]] <= vals[node] ? 2 * node + 1 : 2 * node + 2;
    return node - q;
}

// brute-force.fut:5-22, sequential insert (an equal distance never displaces)
__device__ void ocl_brute_force(const float *qv, int d, const float *pts, int beg, int npts, int *ki, float *kd) {
    for (int i = 0; i < npts; ++i) {
        float dist = ocl_sqdist(qv, pts + (size_t)i * d, d);
        if (dist > kd[kOK - 1]) continue;
        int ref = beg + i;
#pragma unroll
        for (int j = 0; j < kOK; ++j) {
            const float c = kd[j];
            if (dist >= c) continue;
            const int t = ki[j];
            ki[j] = ref;
            kd[j] = dist;
            ref = t;
            dist = c;
        }
    }
}

// brute-force.fut:25-66.  The repeated arg-min extraction visits the leaf's
// points in ascending (distance, position) order, so its first eight are
// the eight smallest such pairs, found here in one pass; they replace the
// list's largest entries from the end while strictly smaller, and the two
// sorted runs are merged taking the back run on ties.
__device__ void ocl_brute_force_par(const float *qv, int d, const float *pts, int beg, int npts, int *ki, float *kd) {
    float cd[kOK];
    int ci[kOK];
#pragma unroll
    for (int j = 0; j < kOK; ++j) { cd[j] = INFINITY; ci[j] = 0x7FFFFFFF; }
    for (int i = 0; i < npts; ++i) {
        float dist = ocl_sqdist(qv, pts + (size_t)i * d, d);
        int idx = i;
        // (dist, idx) enters when below the last pair; positions ascend, so
        // an equal distance goes after the pairs already held
        if (!(dist < cd[kOK - 1] || (dist == cd[kOK - 1] && idx < ci[kOK - 1]))) continue;
#pragma unroll
        for (int j = 0; j < kOK; ++j) {
            const bool below = dist < cd[j] || (dist == cd[j] && idx < ci[j]);
            if (below) {
                const float td = cd[j];
                const int ti = ci[j];
                cd[j] = dist;
                ci[j] = idx;
                dist = td;
                idx = ti;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kOK; ++j) {
        if (!(cd[j] < kd[kOK - 1 - j])) break;
        ki[kOK - 1 - j] = beg + ci[j];
        kd[kOK - 1 - j] = cd[j];
    }
    int oi[kOK];
    float od[kOK];
    int b = 0, e = kOK - 1;
#pragma unroll
    for (int i = 0; i < kOK; ++i) {
        if (kd[b] < kd[e]) { oi[i] = ki[b]; od[i] = kd[b]; ++b; }
        else { oi[i] = ki[e]; od[i] = kd[e]; --e; }
    }
#pragma unroll
    for (int i = 0; i < kOK; ++i) { ki[i] = oi[i]; kd[i] = od[i]; }
}

// findNaturalLeaves (driverKNN.fut:133-152): one thread per query
__global__ void k_ocl_natural(const float *__restrict__ qa, int n_a, int d, const int *dims, const float *vals, int q,
                              const float *__restrict__ leaf_pts, int ppl, int *nat, int *pos, float *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_a) return;
    float qv[kOclMaxD];
    for (int j = 0; j < d; ++j) qv[j] = qa[(size_t)i * d + j];
    const int leaf = ocl_find_leaf(dims, vals, q, qv);
    int ki[kOK];
    float kd[kOK];
#pragma unroll
    for (int k = 0; k < kOK; ++k) { ki[k] = -1; kd[k] = FLT_MAX; }
    ocl_brute_force(qv, d, leaf_pts + (size_t)leaf * ppl * d, leaf * ppl, ppl, ki, kd);
    nat[i] = leaf;
#pragma unroll
    for (int k = 0; k < kOK; ++k) { pos[(size_t)i * kOK + k] = ki[k]; dst[(size_t)i * kOK + k] = kd[k]; }
}

__device__ __forceinline__ int ocl_level(int node) { return 31 - __clz(node + 1); }

// kd-traverse.fut:44-57: the first visited ancestor on the chain of closest
// same-dimension ancestors contributes median - q
__device__ __forceinline__ float ocl_anc_contrib(const int *clanc, const float *vals, float qmi, int stack, int node) {
    int idx = node;
    while (idx >= 0) {
        const int anc = clanc[idx];
        if (anc == -1) return 0.0f;
        if (!((stack >> ocl_level(anc)) & 1)) idx = anc;
        else return vals[anc] - qmi;
    }
    return 0.0f;
}

// kd-traverse.fut:27-116 traverseOnce: from the last scanned leaf, back up
// while the second child was visited or lies beyond the K-th distance, then
// descend near-first to the next leaf (q + 1 when none is left)
__device__ void ocl_traverse_once(int height, int q, const int *dims, const float *vals, const int *clanc,
                                  const float *qv, float wnn, int &leaf, int &stack, float &dist) {
    if (leaf >= q + 1) return;
    int node = leaf + q, count = height, rec = -1, stk = stack;
    float ds = dist;
    while (node != 0 && rec < 0) {
        const int par = (node - 1) >> 1;
        const int scnd = (stk >> count) & 1;
        const float qmd = qv[dims[par]];
        const float cur = vals[par] - qmd;
        const float cur2 = cur * cur;
        const float prv = ocl_anc_contrib(clanc, vals, qmd, stk, par);
        const float prv2 = prv * prv;
        const float t_minus = ds - cur2, t_plus = ds - prv2;
        const float dminus = fabsf(t_minus + prv2);
        const float dplus = fabsf(t_plus + cur2);
        if (scnd) {
            node = par;
            --count;
            ds = dminus;
        } else if (!(dplus < wnn)) {
            node = par;
            --count;
        } else {
            stk |= 1 << count;
            rec = (node % 2 == 0) ? node - 1 : node + 1;
            ds = dplus;
            node = par;
        }
    }
    if (node == 0 && rec == -1) {
        leaf = q + 1;
        stack = stk;
        dist = ds;
        return;
    }
    int n = rec;
    while (n < q) {
        ++count;
        stk &= ~(1 << count);
        n = qv[dims[n]] <= vals[n] ? 2 * n + 1 : 2 * n + 2;
    }
    leaf = n - q;
    stack = stk;
    dist = ds;
}

// exactKnn (driverKNN.fut:209-243) for the first nq queries (row 0), one
// query per lane, from their natural-leaf lists
__global__ void k_ocl_exact(const float *__restrict__ qa, int nq, int d, const int *dims, const float *vals,
                            const int *clanc, int height, const float *__restrict__ leaf_pts, int ppl, const int *nat,
                            int *pos, float *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int q = (1 << (height + 1)) - 1, nleaves = q + 1;
    float qv[kOclMaxD];
    for (int j = 0; j < d; ++j) qv[j] = qa[(size_t)i * d + j];
    int ki[kOK];
    float kd[kOK];
#pragma unroll
    for (int k = 0; k < kOK; ++k) { ki[k] = pos[(size_t)i * kOK + k]; kd[k] = dst[(size_t)i * kOK + k]; }
    int leaf = nat[i], stack = 0;
    float dist = 0.0f;
    // every call either ends (leaf = q + 1) or descends to a leaf right of
    // the last one, so at most nleaves + 1 calls are made
    for (int it = 0; it <= nleaves; ++it) {
        ocl_traverse_once(height, q, dims, vals, clanc, qv, kd[kOK - 1], leaf, stack, dist);
        if (leaf >= nleaves) break;
        ocl_brute_force_par(qv, d, leaf_pts + (size_t)leaf * ppl * d, leaf * ppl, ppl, ki, kd);
    }
#pragma unroll
    for (int k = 0; k < kOK; ++k) { pos[(size_t)i * kOK + k] = ki[k]; dst[(size_t)i * kOK + k] = kd[k]; }
}

// propagateFixK (driverKNN.fut:250-350): one thread per image column walks
// its rows; row y merges, into its natural-leaf list, the leaves of the B
// patches one row below (clamped) the row above's list, without its natural
// leaf and repeats, in list order
__global__ void k_ocl_propagate(const float *__restrict__ qa, int n_rows, int n_cols, int d,
                                const float *__restrict__ leaf_pts, int ppl, const int *__restrict__ indir,
                                const int *__restrict__ orig2leaf, const int *__restrict__ nat, int *pos,
                                float *dst) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n_cols) return;
    int up[kOK];
#pragma unroll
    for (int k = 0; k < kOK; ++k) up[k] = pos[(size_t)x * kOK + k];
    float qv[kOclMaxD];
    for (int y = 1; y < n_rows; ++y) {
        const size_t i = (size_t)y * n_cols + x;
        int ul[kOK], nu = 0;
        const int nl = nat[i];
#pragma unroll
        for (int k = 0; k < kOK; ++k) {
            if (up[k] < 0) continue;
            const int o = indir[up[k]];
            const int oy = o / n_cols, ox = o - oy * n_cols;
            const int cy = oy < n_rows - 1 ? oy + 1 : n_rows - 1;
            const int l = orig2leaf[cy * n_cols + ox];
            bool found = l == nl;
            for (int j = 0; j < nu && !found; ++j) found = ul[j] == l;
            if (!found) ul[nu++] = l;
        }
        int ki[kOK];
        float kd[kOK];
#pragma unroll
        for (int k = 0; k < kOK; ++k) { ki[k] = pos[i * kOK + k]; kd[k] = dst[i * kOK + k]; }
        if (nu > 0) {
            for (int j = 0; j < d; ++j) qv[j] = qa[i * d + j];
            for (int j = 0; j < nu; ++j)
                ocl_brute_force_par(qv, d, leaf_pts + (size_t)ul[j] * ppl * d, ul[j] * ppl, ppl, ki, kd);
        }
#pragma unroll
        for (int k = 0; k < kOK; ++k) {
            pos[i * kOK + k] = ki[k];
            dst[i * kOK + k] = kd[k];
            up[k] = ki[k];
        }
    }
}

// positions -> original indices (driverKNN.fut:409) and selectBestNN
// (driverKNN.fut:93-125): the first strict minimum of the exact full-patch
// distance over the eight candidates (an f32.sum of integers, exact below 2^24)
__global__ void k_ocl_best(const int *__restrict__ pos, const int *__restrict__ indir, int n_a, const uint8_t *img_a,
                           int wa, int nca, const uint8_t *img_b, int wb, int ncb, int p, int *knn, int *nn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_a) return;
    int cand[kOK];
#pragma unroll
    for (int k = 0; k < kOK; ++k) {
        const int ps = pos[(size_t)i * kOK + k];
        cand[k] = ps >= 0 ? indir[ps] : -1;
        if (knn) knn[(size_t)i * kOK + k] = cand[k];
    }
    if (!nn) return;
    const int ya = i / nca, xa = i - ya * nca;
    const uint8_t *a0 = img_a + ((size_t)ya * wa + xa) * 3;
    int best = -1;
    long long bd = 0;
#pragma unroll 1
    for (int k = 0; k < kOK; ++k) {
        const int j = cand[k];
        if (j < 0) continue;
        const int yb = j / ncb, xb = j - yb * ncb;
        const uint8_t *b0 = img_b + ((size_t)yb * wb + xb) * 3;
        long long s = 0;
        for (int r = 0; r < p; ++r) {
            const uint8_t *ar = a0 + (size_t)r * wa * 3, *br = b0 + (size_t)r * wb * 3;
            int rs = 0;
            for (int f = 0; f < 3 * p; ++f) {
                const int t = (int)br[f] - (int)ar[f];
                rs += t * t;
            }
            s += rs;
        }
        if (best < 0 || s < bd) { best = j; bd = s; }
    }
    nn[i] = best;
}

}  // namespace

void ocl_tree_shape(int m, int defppl, int &height, int &q, int &ppl, int &mp) {
    const int dnl = (m + defppl - 1) / defppl;
    int hp1 = 0;
    if (dnl > 1) {
        int r = 0;
        for (int t = dnl; t > 1; t >>= 1) ++r;
        const long long down = (long long)dnl - (1LL << r), up = (1LL << (r + 1)) - dnl;
        hp1 = down <= up ? r : r + 1;
    }
    if (hp1 <= 0) {
        height = -1;
        q = 0;
        ppl = m;
        mp = m;
        return;
    }
    height = hp1 - 1;
    const int nl = 1 << (height + 1);
    q = nl - 1;
    ppl = (m + nl - 1) / nl;
    mp = nl * ppl;
}

size_t ocl_sort_bytes(int mp) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                    (int *)nullptr, (int *)nullptr, (unsigned)mp, 0u, 64u);
    return bytes;
}

hipError_t ocl_build(const float *X, int m, int d, int defppl, const OclTree &t, void *sort_tmp, size_t sort_bytes,
                     hipStream_t st) {
    hipError_t e;
    const unsigned b = 256;
    hipLaunchKernelGGL(k_ocl_init, dim3((unsigned)((std::max(t.mp, d) + 255) / 256)), dim3(b), 0, st, t.lub, d,
                       t.indir, t.mp);
    hipLaunchKernelGGL(k_ocl_bounds, dim3(512), dim3(b), 0, st, X, m, d, t.lub);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    int *ind = t.indir, *ind2 = t.indir2;
    for (int lev = 0; lev <= t.height; ++lev) {
        const unsigned nn = (unsigned)((1 << lev) + 255) / 256;
        hipLaunchKernelGGL(k_ocl_dims, dim3(nn), dim3(b), 0, st, t.lub, d, lev, t.dims, t.vals, t.clanc);
        hipLaunchKernelGGL(k_ocl_keys, dim3((unsigned)((t.mp + 255) / 256)), dim3(b), 0, st, X, m, d, t.mp, lev,
                           t.dims, ind, t.keys);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        size_t bytes = sort_bytes;
        // stable: equal (node, key) pairs keep their order
        e = rocprim::radix_sort_pairs(sort_tmp, bytes, t.keys, t.keys2, ind, ind2, (unsigned)t.mp, 0u,
                                      (unsigned)(32 + lev), st);
        if (e != hipSuccess) return e;
        std::swap(ind, ind2);
        hipLaunchKernelGGL(k_ocl_medians, dim3(nn), dim3(b), 0, st, X, m, d, t.mp, lev, t.dims, ind, t.vals);
    }
    if (ind != t.indir) {
        e = hipMemcpyAsync(t.indir, ind, sizeof(int) * (size_t)t.mp, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_ocl_leaves, dim3((unsigned)(((long long)t.mp * d + 255) / 256)), dim3(b), 0, st, X, m, d,
                       t.mp, t.ppl, t.indir, t.leaf_pts, t.orig2leaf);
    return hipGetLastError();
}

hipError_t ocl_queries(const OclTree &t, const float *qa, int n_rows, int n_cols, int d, int *nat, int *pos,
                       float *dst, hipStream_t st) {
    const int n_a = n_rows * n_cols;
    const unsigned b = 256;
    hipLaunchKernelGGL(k_ocl_natural, dim3((unsigned)((n_a + 255) / 256)), dim3(b), 0, st, qa, n_a, d, t.dims, t.vals,
                       t.q, t.leaf_pts, t.ppl, nat, pos, dst);
    hipLaunchKernelGGL(k_ocl_exact, dim3((unsigned)((n_cols + 63) / 64)), dim3(64), 0, st, qa, std::min(n_cols, n_a),
                       d, t.dims, t.vals, t.clanc, t.height, t.leaf_pts, t.ppl, nat, pos, dst);
    if (n_rows > 1)
        hipLaunchKernelGGL(k_ocl_propagate, dim3((unsigned)((n_cols + 63) / 64)), dim3(64), 0, st, qa, n_rows, n_cols,
                           d, t.leaf_pts, t.ppl, t.indir, t.orig2leaf, nat, pos, dst);
    return hipGetLastError();
}

hipError_t ocl_best(const int *pos, const int *indir, int n_a, const uint8_t *img_a, int wa, int nca,
                    const uint8_t *img_b, int wb, int ncb, int p, int *knn, int *nn, hipStream_t st) {
    hipLaunchKernelGGL(k_ocl_best, dim3((unsigned)((n_a + 255) / 256)), dim3(256), 0, st, pos, indir, n_a, img_a, wa,
                       nca, img_b, wb, ncb, p, knn, nn);
    return hipGetLastError();
}

}  // namespace annf
