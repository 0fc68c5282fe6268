// Exact nearest-neighbour field on full uint8 patches (the evaluation
// reference for every variant): squared L2 distances are exact integers,
// |a'|^2 + |b'|^2 - 2 a'.b' of the centred bytes x' = x - 128; the k nearest
// B patches of each sampled A patch are kept in the order (distance, index).
// Every value is below 2^25, so this equals an FP64 brute force bit for bit.
//
// The dot products run on v_mfma_i32_32x32x32_i8: x - 128 is exactly a
// signed byte (x ^ 0x80) and the accumulation is exact in int32 (at most 512
// products of magnitude <= 16384).  Patches are rows of DQ = 32 ceil(D / 32)
// centred bytes (zero padded); a block holds 128 sampled queries, 32 per
// wavefront with their bytes in registers (lane l: query l & 31, bytes
// 16 (l >> 5) .. + 15 of every 32-byte k-step), and walks one contiguous
// range of B patches in LDS-staged tiles of 32 (double buffered).  Keys
// (|b'|^2 + OFF - 2 a'.b') << PB | position within a 2^PB-patch segment are
// kept per lane in a v_med3 network; at a segment's end the two lanes of a
// query merge theirs and the k best become (d2 << 32 | B index) keys of the
// block's list; per-split lists are merged per query at the end.
#include "annf_internal.h"
#include "annf_lists.h"

#include <type_traits>

namespace annf {

namespace {

typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

inline unsigned blocks_of(long long n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

// Centred patch rows: row i holds patch idx[i] (or i) of the image, byte f =
// (r*p + c)*3 + ch of the window as x ^ 0x80, zero from D up to dq (a
// multiple of 32) -- one dword per thread.
__global__ void k_patch_matrix(const uint8_t *__restrict__ img, int w, int p, int ncols, const int *__restrict__ idx,
                               long long n, int dq, uint8_t *__restrict__ out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = n * (dq / 4);
    if (t >= total) return;
    const long long row = t / (dq / 4);
    const int wd = (int)(t - row * (dq / 4));
    const long long i = idx ? idx[row] : row;
    const long long y = i / ncols, x = i - y * ncols;
    const int D = p * p * 3, rl = 3 * p;
    uint32_t v = 0;
    for (int b = 0; b < 4; ++b) {
        const int f = wd * 4 + b;
        if (f < D) {
            const int r = f / rl, u = f - r * rl;
            v |= (uint32_t)(img[((y + r) * w + x) * 3 + u] ^ 0x80u) << (8 * b);
        }
    }
    reinterpret_cast<uint32_t *>(out)[t] = v;
}

// |x'|^2 of each centred row
__global__ void k_patch_norms(const uint8_t *__restrict__ m, long long n, int dq, uint32_t *__restrict__ nrm) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int *r = reinterpret_cast<const int *>(m + i * dq);
    int s = 0;
    for (int w = 0; w < dq / 4; ++w) s = __builtin_amdgcn_sdot4(r[w], r[w], s, false);
    nrm[i] = (uint32_t)s;
}

constexpr int kEQ = 128;          // queries per block
constexpr int kETP = 32;          // B patches per tile
constexpr int kEThreads = 256;

// Block (blockIdx.x, split blockIdx.y): queries 128 x .. + 127 (rows of am)
// against B patches [split * per_split, + per_split).  part[(q * splits +
// split) * k ...] receives the block's k best keys ((d2 << 32 | index) + 1,
// ascending; ~0 past the candidates).
template <int KS, int LM>
__global__ __launch_bounds__(kEThreads, 2) void k_exact_mfma(const uint8_t *__restrict__ am,
                                                           const uint32_t *__restrict__ an, int nq,
                                                           const uint8_t *__restrict__ bm,
                                                           const uint32_t *__restrict__ bn, long long nb, int k,
                                                           long long per_split, int pb, uint32_t off,
                                                           uint64_t *__restrict__ part) {
    constexpr int DQ = 32 * KS;
    constexpr int CH = DQ / 16;                           // 16-byte chunks per patch
    constexpr int SETS = kEThreads / kETP;                // staging threads per patch
    constexpr int CPT = (CH + SETS - 1) / SETS;
    __shared__ __align__(16) uint8_t bt[2][CH * kETP * 16];   // [chunk][patch][16 bytes]
    __shared__ __align__(16) uint32_t base[2][kETP];

    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int h = lane >> 5, qc = lane & 31;
    const int qi = blockIdx.x * kEQ + 32 * wid + qc;
    const bool qv = qi < nq;
    const bool wave_live = blockIdx.x * kEQ + 32 * wid < nq;
    const int split = blockIdx.y;
    const long long b0 = (long long)split * per_split;
    const long long b1 = b0 + per_split < nb ? b0 + per_split : nb;
    const int seg = 1 << pb;
    const int mul = -(2 << pb);

    intx4 af[KS];
    uint32_t na = 0;
    {
        const int q = qv ? qi : 0;
        na = an[q];
        const intx4 *src = reinterpret_cast<const intx4 *>(am + (size_t)q * DQ + 16 * h);
#pragma unroll
        for (int s = 0; s < KS; ++s) af[s] = src[2 * s];
    }

    const int sp = tid / SETS, sc = tid % SETS;
    // tile loads run two tiles ahead: register set A or B holds the tile
    // after the one in LDS, the other receives the one after that
    auto fetch = [&](u32x4 (&r)[CPT], uint32_t &rb, long long t) {
        const long long j = t + sp;
        const bool v = j < b1;
        const u32x4 *src = reinterpret_cast<const u32x4 *>(bm + (size_t)(v ? j : b0) * DQ);
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int c = sc + SETS * i;
            r[i] = (v && c < CH) ? src[c] : u32x4{0u, 0u, 0u, 0u};
        }
        if (sc == 0) rb = v ? (((bn[j] + off) << pb) | (uint32_t)((j - b0) & (seg - 1))) : 0xFFFFFFFFu;
    };
    auto store = [&](int buf, const u32x4 (&r)[CPT], uint32_t rb) {
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int c = sc + SETS * i;
            if (c < CH) *reinterpret_cast<u32x4 *>(bt[buf] + (c * kETP + sp) * 16) = r[i];
        }
        if (sc == 0) base[buf][sp] = rb;
    };

    uint32_t l32[LM];
    uint64_t l64[LM];
#pragma unroll
    for (int s = 0; s < LM; ++s) {
        l32[s] = s < LM - k ? 0u : 0xFFFFFFFFu;
        l64[s] = s < LM - k ? 0ull : ~0ull;
    }
    u32x4 rA[CPT], rB[CPT];
    uint32_t bA = 0xFFFFFFFFu, bB = 0xFFFFFFFFu;
    long long t0 = b0;
    if (b0 < b1) {
        fetch(rA, bA, b0);
        store(0, rA, bA);
        if (b0 + kETP < b1) fetch(rA, bA, b0 + kETP);
    }
    __syncthreads();
    int cur = 0;
    // tile t0 is in LDS buffer cur, tile t0 + 32 in rs; tile t0 + 64 goes into rf
    auto step = [&](u32x4 (&rs)[CPT], uint32_t &bs, u32x4 (&rf)[CPT], uint32_t &bf) {
        const long long tt0 = t0;
        const bool seg_end = tt0 + kETP >= b1 || (((tt0 - b0) + kETP) & (seg - 1)) == 0;
        if (tt0 + 2 * kETP < b1) fetch(rf, bf, tt0 + 2 * kETP);
        if (wave_live) {
            intx16 acc = {}, acc1 = {};
            const uint8_t *br = bt[cur] + (h * kETP + qc) * 16;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const intx4 bv = *reinterpret_cast<const intx4 *>(br + (2 * s) * kETP * 16);
                if (s & 1) acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(bv, af[s], acc1, 0, 0, 0);
                else acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(bv, af[s], acc, 0, 0, 0);
            }
            acc += acc1;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const u32x4 bq = *reinterpret_cast<const u32x4 *>(&base[cur][8 * j + 4 * h]);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    list_insert<LM>(l32, bq[i] + (uint32_t)__mul24(acc[4 * j + i], mul));
            }
            if (seg_end) {
                // the segment's k best of this query: merge the two lanes'
                // lists, then as (d2, B index) keys into the running list
                const long long segbase = b0 + ((tt0 - b0) & ~(long long)(seg - 1));
                uint32_t other[LM];
#pragma unroll
                for (int s = 0; s < LM; ++s) other[s] = (uint32_t)__shfl_xor((int)l32[s], 32, 64);
#pragma unroll
                for (int s = 0; s < LM; ++s)
                    if (s >= LM - k) list_insert<LM>(l32, other[s]);
#pragma unroll
                for (int s = 0; s < LM; ++s) {
                    if (s < LM - k) continue;
                    const uint32_t kk = l32[s];
                    if (kk != 0xFFFFFFFFu) {
                        const uint64_t key = (((uint64_t)((kk >> pb) - off + na) << 32) |
                                              (uint64_t)(segbase + (kk & (uint32_t)(seg - 1)))) + 1ull;
                        if (key < l64[LM - 1]) list_insert64<LM>(l64, key);
                    }
                    l32[s] = 0xFFFFFFFFu;
                }
            }
        }
        if (tt0 + kETP < b1) store(cur ^ 1, rs, bs);
        __syncthreads();
        cur ^= 1;
        t0 += kETP;
    };
    for (;;) {
        if (t0 >= b1) break;
        step(rA, bA, rB, bB);
        if (t0 >= b1) break;
        step(rB, bB, rA, bA);
    }
    if (qv && h == 0) {
        uint64_t *o = part + ((size_t)qi * gridDim.y + split) * k - (LM - k);
#pragma unroll
        for (int s = 0; s < LM; ++s)
            if (s >= LM - k) o[s] = l64[s];
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
        if (key < lst[KS - 1]) list_insert64<KS>(lst, key);
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
hipError_t with_ks_k(int ks, int k, Fn &&fn) {
    auto pick = [&](auto kk) -> hipError_t {
        if (k <= 5) return fn(kk, std::integral_constant<int, 5>{});
        if (k <= 8) return fn(kk, std::integral_constant<int, 8>{});
        return fn(kk, std::integral_constant<int, 32>{});
    };
    switch (ks) {
    case 1: return pick(std::integral_constant<int, 1>{});
    case 2: return pick(std::integral_constant<int, 2>{});
    case 3: return pick(std::integral_constant<int, 3>{});
    case 4: return pick(std::integral_constant<int, 4>{});
    case 5: return pick(std::integral_constant<int, 5>{});
    case 6: return pick(std::integral_constant<int, 6>{});
    case 8: return pick(std::integral_constant<int, 8>{});
    case 10: return pick(std::integral_constant<int, 10>{});
    case 12: return pick(std::integral_constant<int, 12>{});
    case 14: return pick(std::integral_constant<int, 14>{});
    case 16: return pick(std::integral_constant<int, 16>{});
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int exact_dq(int p) {
    const int D = p * p * 3;
    const int ks = (D + 31) / 32;
    // the k-step counts instantiated: 1-6, then even counts up to 16
    const int ksi = ks <= 6 ? ks : (ks + 1) / 2 * 2;
    return ksi <= 16 ? 32 * ksi : 0;
}

hipError_t patch_matrix(const uint8_t *img, int h, int w, int p, const int *idx, long long n, uint8_t *out, int dq,
                        hipStream_t st) {
    const long long total = n * (dq / 4);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_patch_matrix, dim3(blocks_of(total, 256)), dim3(256), 0, st, img, w, p, w - p + 1, idx, n,
                       dq, out);
    return hipGetLastError();
}

hipError_t patch_norms(const uint8_t *m, long long n, int dq, uint32_t *nrm, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_patch_norms, dim3(blocks_of(n, 256)), dim3(256), 0, st, m, n, dq, nrm);
    return hipGetLastError();
}

int exact_splits(long long nb) {
    long long s = (nb + 32767) / 32768;
    return (int)(s < 1 ? 1 : (s > 128 ? 128 : s));
}

hipError_t exact_sample(const uint8_t *am, const uint32_t *an, int nq, const uint8_t *bm, const uint32_t *bn,
                        long long nb, int p, int dq, int k, uint64_t *part, int *idx, int *d2, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    const int D = p * p * 3;
    const int splits = exact_splits(nb);
    const long long per = (nb + splits - 1) / splits;
    // OFF = 16384 D keeps |b'|^2 + OFF - 2 a'.b' in [0, 65536 D]; PB: the
    // position bits that leave room for it in 32 bits
    const uint32_t off = 16384u * (uint32_t)D;
    const uint64_t vmax = 65536ull * (uint64_t)D;
    int pb = 7;
    while (pb > 0 && ((vmax << pb) | ((1ull << pb) - 1)) >= 0xFFFFFFFFull) --pb;
    if (pb < 5) return hipErrorInvalidValue;
    return with_ks_k(dq / 32, k, [&](auto ks, auto lm) -> hipError_t {
        constexpr int KS = decltype(ks)::value;
        constexpr int LM = decltype(lm)::value;
        hipLaunchKernelGGL((k_exact_mfma<KS, LM>), dim3(blocks_of(nq, kEQ), splits), dim3(kEThreads), 0, st, am, an,
                           nq, bm, bn, nb, k, per, pb, off, part);
        hipLaunchKernelGGL((k_exact_merge<LM>), dim3(blocks_of(nq, 256)), dim3(256), 0, st, part, nq, splits, k, idx,
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
