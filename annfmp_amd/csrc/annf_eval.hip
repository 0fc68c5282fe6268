// Exact nearest-neighbour field on full uint8 patches (the evaluation
// reference for every variant): squared L2 distances are exact integers,
// |a|^2 + |b|^2 - 2 a.b with 4-byte dot products (v_dot4_u32_u8); the k
// nearest B patches of each sampled A patch are kept in the order
// (distance, index).  Every value is below 2^25, so this equals an FP64
// brute force bit for bit.
#include "annf_internal.h"

#include <type_traits>

namespace annf {

namespace {

inline unsigned blocks_of(long long n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

// Patch rows of an image as a zero-padded byte matrix: row i holds the
// p x p x 3 window of patch i in feature order f = (r*p + c)*3 + ch, padded
// to dpad_b bytes (a multiple of 4).
__global__ void k_patch_matrix(const uint8_t *__restrict__ img, int w, int p, int ncols, long long n, int dpad_b,
                               uint8_t *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = n * (dpad_b / 4);
    if (t >= total) return;
    const long long i = t / (dpad_b / 4);
    const int wd = (int)(t - i * (dpad_b / 4));
    const long long y = i / ncols, x = i - y * ncols;
    const int D = p * p * 3, rl = 3 * p;
    uint32_t v = 0;
    for (int b = 0; b < 4; ++b) {
        const int f = wd * 4 + b;
        if (f < D) {
            const int r = f / rl, u = f - r * rl;
            v |= (uint32_t)img[((y + r) * w + x) * 3 + u] << (8 * b);
        }
    }
    reinterpret_cast<uint32_t *>(out)[t] = v;
}

__global__ void k_patch_norms(const uint8_t *__restrict__ m, long long n, int dpad_b, uint32_t *__restrict__ nrm) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t *r = reinterpret_cast<const uint32_t *>(m + i * dpad_b);
    uint32_t s = 0;
    for (int w = 0; w < dpad_b / 4; ++w) s = __builtin_amdgcn_udot4(r[w], r[w], s, false);
    nrm[i] = s;
}

constexpr int kTile = 64;   // B patches staged in LDS per step

// One thread per sampled query; the block walks a contiguous range of B
// patches in tiles staged in LDS.  Keys (d2 << 32 | index) + 1 are kept in a
// sorted list whose live entries sit in the top k slots (front slots hold 0,
// below every key), so no slot index depends on k.
template <int MAXW, int KS>
__global__ __launch_bounds__(256) void k_exact_sample(const uint8_t *__restrict__ am, const uint32_t *__restrict__ an,
                                                      const int *__restrict__ qidx, int nq,
                                                      const uint8_t *__restrict__ bm,
                                                      const uint32_t *__restrict__ bn, long long nb, int nw, int k,
                                                      long long per_split, uint64_t *__restrict__ part) {
    __shared__ uint32_t tile[kTile * MAXW];
    __shared__ uint32_t tnorm[kTile];
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    const int split = blockIdx.y;
    const bool live = qi < nq;
    uint32_t qa[MAXW];
    uint32_t na = 0;
    if (live) {
        const long long a = qidx[qi];
        const uint32_t *r = reinterpret_cast<const uint32_t *>(am + a * (long long)nw * 4);
#pragma unroll
        for (int w = 0; w < MAXW; ++w) qa[w] = w < nw ? r[w] : 0u;
        na = an[a];
    } else {
#pragma unroll
        for (int w = 0; w < MAXW; ++w) qa[w] = 0u;
    }
    uint64_t lst[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) lst[s] = (s < KS - k) ? 0ull : ~0ull;
    const long long b0 = (long long)split * per_split;
    const long long b1 = b0 + per_split < nb ? b0 + per_split : nb;
    for (long long tb = b0; tb < b1; tb += kTile) {
        const int cnt = (int)(b1 - tb < kTile ? b1 - tb : kTile);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * nw; e += blockDim.x) {
            const int j = e / nw, w = e - j * nw;
            tile[j * MAXW + w] = reinterpret_cast<const uint32_t *>(bm + (tb + j) * (long long)nw * 4)[w];
        }
        for (int j = threadIdx.x; j < cnt; j += blockDim.x) tnorm[j] = bn[tb + j];
        __syncthreads();
        if (!live) continue;
        for (int j = 0; j < cnt; ++j) {
            const uint32_t *bt = tile + j * MAXW;
            uint32_t dot = 0;
#pragma unroll
            for (int w = 0; w < MAXW; ++w) dot = __builtin_amdgcn_udot4(qa[w], bt[w], dot, false);
            const uint32_t d2 = na + tnorm[j] - 2u * dot;
            const uint64_t key = (((uint64_t)d2 << 32) | (uint64_t)(tb + j)) + 1ull;
            if (key >= lst[KS - 1]) continue;
            int r = 0;
#pragma unroll
            for (int s = 0; s < KS - 1; ++s) r += lst[s] < key ? 1 : 0;
#pragma unroll
            for (int s = KS - 1; s >= 1; --s)
                if (s > r) lst[s] = lst[s - 1];
#pragma unroll
            for (int s = 0; s < KS; ++s)
                if (s == r) lst[s] = key;
        }
    }
    if (live) {
        uint64_t *o = part + ((size_t)qi * gridDim.y + split) * k - (KS - k);
#pragma unroll
        for (int s = 0; s < KS; ++s)
            if (s >= KS - k) o[s] = lst[s];
    }
}

// Merges the per-split lists of each query (one thread per query).
template <int KS>
__global__ void k_exact_merge(const uint64_t *__restrict__ part, int nq, int splits, int k, int *__restrict__ idx,
                              int *__restrict__ d2) {
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    uint64_t lst[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) lst[s] = (s < KS - k) ? 0ull : ~0ull;
    const uint64_t *p = part + (size_t)qi * splits * k;
    for (int e = 0; e < splits * k; ++e) {
        const uint64_t key = p[e];
        if (key >= lst[KS - 1]) continue;
        int r = 0;
#pragma unroll
        for (int s = 0; s < KS - 1; ++s) r += lst[s] < key ? 1 : 0;
#pragma unroll
        for (int s = KS - 1; s >= 1; --s)
            if (s > r) lst[s] = lst[s - 1];
#pragma unroll
        for (int s = 0; s < KS; ++s)
            if (s == r) lst[s] = key;
    }
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s >= KS - k) {
            const uint64_t key = lst[s] - 1ull;
            idx[(size_t)qi * k + s - (KS - k)] = (int)(uint32_t)key;
            d2[(size_t)qi * k + s - (KS - k)] = (int)(key >> 32);
        }
    }
}

template <typename Fn>
hipError_t with_words(int nw, int k, Fn &&fn) {
    auto pick_k = [&](auto mw) {
        if (k <= 8) return fn(mw, std::integral_constant<int, 8>{});
        return fn(mw, std::integral_constant<int, 32>{});
    };
    if (nw <= 48) return pick_k(std::integral_constant<int, 48>{});
    if (nw <= 96) return pick_k(std::integral_constant<int, 96>{});
    if (nw <= 128) return pick_k(std::integral_constant<int, 128>{});
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t patch_matrix(const uint8_t *img, int h, int w, int p, uint8_t *out, int dpad_b, hipStream_t st) {
    const long long n = (long long)(h - p + 1) * (w - p + 1);
    const long long total = n * (dpad_b / 4);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_patch_matrix, dim3(blocks_of(total, 256)), dim3(256), 0, st, img, w, p, w - p + 1, n,
                       dpad_b, out);
    return hipGetLastError();
}

hipError_t patch_norms(const uint8_t *m, long long n, int dpad_b, uint32_t *nrm, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_patch_norms, dim3(blocks_of(n, 256)), dim3(256), 0, st, m, n, dpad_b, nrm);
    return hipGetLastError();
}

int exact_splits(long long nb) {
    long long s = (nb + 32767) / 32768;
    return (int)(s < 1 ? 1 : (s > 128 ? 128 : s));
}

hipError_t exact_sample(const uint8_t *am, const uint32_t *an, const int *qidx, int nq, const uint8_t *bm,
                        const uint32_t *bn, long long nb, int dpad_b, int k, uint64_t *part, int *idx, int *d2,
                        hipStream_t st) {
    if (nq == 0) return hipSuccess;
    const int nw = dpad_b / 4;
    const int splits = exact_splits(nb);
    const long long per = (nb + splits - 1) / splits;
    return with_words(nw, k, [&](auto mw, auto ks) {
        constexpr int MAXW = decltype(mw)::value;
        constexpr int KS = decltype(ks)::value;
        hipLaunchKernelGGL((k_exact_sample<MAXW, KS>), dim3(blocks_of(nq, 256), splits), dim3(256), 0, st, am, an,
                           qidx, nq, bm, bn, nb, nw, k, per, part);
        hipLaunchKernelGGL((k_exact_merge<KS>), dim3(blocks_of(nq, 256)), dim3(256), 0, st, part, nq, splits, k, idx,
                           d2);
        return hipGetLastError();
    });
}

}  // namespace annf

// ---------------------------------------------------------------------------
// PCA covariance in FP64 (scikit-learn's covariance_eigh form, _pca.py 1.7.2:
// C = X^T X; C -= (n * mean_col) * mean_row; C /= n - 1) of sampled patches.
// On uint8 patches X^T X and the column sums are exact integers (uint64
// accumulation, any order), so only the three FP64 operations per entry
// round, in the library's order: the result equals the CPU computation bit
// for bit.
// ---------------------------------------------------------------------------
namespace annf {
namespace {

// XT[f][s] (D x ns, ns padded to a multiple of 4 with zeros) of the sampled
// patches: samples [0, na) from image a (indices ia), then nb from image b.
__global__ void k_gather_t(PatchGeom ga, const int *__restrict__ ia, int na, PatchGeom gb,
                           const int *__restrict__ ib, int nb, int D, int ns_pad, uint8_t *__restrict__ XT) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ns_pad * D) return;
    const int f = (int)(t / ns_pad), s = (int)(t - (long long)f * ns_pad);
    uint8_t v = 0;
    if (s < na + nb) {
        const PatchGeom &g = s < na ? ga : gb;
        const long long i = s < na ? ia[s] : ib[s - na];
        const long long y = i / g.ncols, x = i - y * g.ncols;
        const int r = f / g.rowlen, u = f - r * g.rowlen;
        v = g.base[(y * g.width + x) * 3 + r * g.rstride + u];
    }
    XT[(size_t)f * ns_pad + s] = v;
}

// column sums S[f] = sum_s x[s][f], exact
__global__ __launch_bounds__(256) void k_feature_sums(const uint8_t *__restrict__ XT, int ns_pad,
                                                      unsigned long long *__restrict__ S) {
    __shared__ unsigned long long part[256];
    const int f = blockIdx.x;
    const uint32_t *row = reinterpret_cast<const uint32_t *>(XT + (size_t)f * ns_pad);
    unsigned long long acc = 0;
    for (int w = threadIdx.x; w < ns_pad / 4; w += 256) acc += __builtin_amdgcn_udot4(row[w], 0x01010101u, 0u, false);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) S[f] = part[0];
}

// Gram G[i][j] = sum_s x[s][i] x[s][j], exact: 32 x 32 tiles, 256-sample
// chunks staged in LDS, dot4 per 4 samples into a per-chunk uint32 (at most
// 256 * 65025 < 2^32), summed in uint64.
__global__ __launch_bounds__(1024) void k_gram(const uint8_t *__restrict__ XT, int ns_pad, int D,
                                               unsigned long long *__restrict__ G) {
    __shared__ uint32_t ti[32][65], tj[32][65];
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    const int ii = threadIdx.x >> 5, jj = threadIdx.x & 31;
    unsigned long long acc = 0;
    const int nw = ns_pad / 4;
    for (int w0 = 0; w0 < nw; w0 += 64) {
        const int cw = min(64, nw - w0);
        for (int e = threadIdx.x; e < 32 * 64; e += 1024) {
            const int r = e >> 6, w = e & 63;
            const bool in = w < cw;
            ti[r][w] = (in && i0 + r < D) ? reinterpret_cast<const uint32_t *>(XT + (size_t)(i0 + r) * ns_pad)[w0 + w]
                                          : 0u;
            tj[r][w] = (in && j0 + r < D) ? reinterpret_cast<const uint32_t *>(XT + (size_t)(j0 + r) * ns_pad)[w0 + w]
                                          : 0u;
        }
        __syncthreads();
        uint32_t part = 0;
        for (int w = 0; w < cw; ++w) part = __builtin_amdgcn_udot4(ti[ii][w], tj[jj][w], part, false);
        acc += part;
        __syncthreads();
    }
    if (i0 + ii < D && j0 + jj < D) G[(size_t)(i0 + ii) * D + j0 + jj] = acc;
}

__global__ void k_cov_final(const unsigned long long *__restrict__ S, const unsigned long long *__restrict__ G, int D,
                            long long n, double *__restrict__ mean, double *__restrict__ C) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < D) mean[t] = (double)S[t] / (double)n;
    if (t >= (long long)D * D) return;
    const int i = (int)(t / D), j = (int)(t - (long long)i * D);
    const double mi = (double)S[i] / (double)n, mj = (double)S[j] / (double)n;
    const double sub = ((double)n * mi) * mj;
    C[t] = ((double)G[t] - sub) / (double)(n - 1);
}

}  // namespace

size_t pca_cov_scratch(int D, int ns) {
    const size_t ns_pad = (size_t)(ns + 3) / 4 * 4;
    return (size_t)D * ns_pad + (size_t)D * sizeof(unsigned long long) + (size_t)D * D * sizeof(unsigned long long) +
           256;
}

hipError_t pca_cov(const PatchGeom &ga, const int *ia, int na, const PatchGeom &gb, const int *ib, int nb, int D,
                   uint8_t *scratch, double *mean, double *C, hipStream_t st) {
    const int ns = na + nb;
    const int ns_pad = (ns + 3) / 4 * 4;
    uint8_t *XT = scratch;
    size_t off = ((size_t)D * ns_pad + 255) / 256 * 256;
    unsigned long long *S = reinterpret_cast<unsigned long long *>(scratch + off);
    unsigned long long *G = S + D;
    const long long tot = (long long)ns_pad * D;
    hipLaunchKernelGGL(k_gather_t, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ga, ia, na, gb, ib, nb, D,
                       ns_pad, XT);
    hipLaunchKernelGGL(k_feature_sums, dim3(D), dim3(256), 0, st, XT, ns_pad, S);
    hipLaunchKernelGGL(k_gram, dim3((D + 31) / 32, (D + 31) / 32), dim3(1024), 0, st, XT, ns_pad, D, G);
    const long long dd = (long long)D * D;
    hipLaunchKernelGGL(k_cov_final, dim3((unsigned)((dd + 255) / 256)), dim3(256), 0, st, S, G, D, (long long)ns,
                       mean, C);
    return hipGetLastError();
}

}  // namespace annf

namespace annf {

// Components from the eigenvectors of the covariance (columns of V, D x D
// column-major, eigenvalues w ascending, as LAPACK's and rocSOLVER's syevd
// return them): component c is eigenvector D-1-c (descending eigenvalue,
// scikit-learn's flip), multiplied by the sign of its first entry of largest
// magnitude (svd_flip with u_based_decision=False, utils/extmath.py), and its
// explained variance max(w, 0).  One block per component.
__global__ __launch_bounds__(256) void k_pca_components(const double *__restrict__ V, const double *__restrict__ w,
                                                        int D, int d, double *__restrict__ comp,
                                                        double *__restrict__ var) {
    __shared__ double sv[256];
    __shared__ int si[256];
    const int c = blockIdx.x, col = D - 1 - c;
    const double *v = V + (size_t)col * D;
    double best = -1.0;
    int bi = D;
    for (int f = threadIdx.x; f < D; f += 256) {
        const double a = fabs(v[f]);
        if (a > best) { best = a; bi = f; }   // f ascending: keeps the first largest
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double ob = sv[threadIdx.x + s];
            const int oi = si[threadIdx.x + s];
            if (ob > sv[threadIdx.x] || (ob == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = ob;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const double pv = v[si[0]];
    const double sgn = pv > 0.0 ? 1.0 : (pv < 0.0 ? -1.0 : 0.0);
    for (int f = threadIdx.x; f < D; f += 256) comp[(size_t)c * D + f] = v[f] * sgn;
    if (var && threadIdx.x == 0) var[c] = w[col] > 0.0 ? w[col] : 0.0;
}

hipError_t pca_components(const double *V, const double *w, int D, int d, double *comp, double *var, hipStream_t st) {
    hipLaunchKernelGGL(k_pca_components, dim3(d), dim3(256), 0, st, V, w, D, d, comp, var);
    return hipGetLastError();
}

}  // namespace annf
