// The learned patch encoder: an MLP D -> h1 -> h2 -> d with GELU (tanh
// form) after the two hidden layers, operands in BF16 on MFMA with FP32
// accumulation.  Training kernels (GEMM, bias + activation, triplet margin
// loss, backward pieces, Adam) and the fused inference kernel that reads
// patches straight from the image and writes the d-dimensional embedding.
//
// Input scaling: a patch byte x enters the network as (x - 128) / 128, which
// BF16 holds exactly.  Parameters are one FP32 buffer:
//   W1 (h1 x D), b1 (h1), W2 (h2 x h1), b2 (h2), W3 (d x h2), b3 (d).
#include "annf_internal.h"
#include "annf_patch.h"

namespace annf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr float kG0 = 1.5957691216057308f;   // 2 sqrt(2 / pi)
constexpr float kG1 = 0.07135481627f;        // 2 sqrt(2 / pi) * 0.044715

inline unsigned nblk(long long n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// written as x * sigmoid(2u) with the hardware exp2 and reciprocal (about
// 1 ulp each; the network rounds activations to BF16 anyway).
__device__ __forceinline__ float sigmoid_f(float t) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-t * 1.4426950408889634f));
}

__device__ __forceinline__ float gelu_f(float x) {
    const float t = x * (kG0 + kG1 * (x * x));
    return x * sigmoid_f(t);
}

// The same GELU for the fused inference kernel, with fused multiply-adds
// and the exp2 scale folded into the polynomial.
__device__ __forceinline__ float gelu_fast(float x) {
    constexpr float kL = 1.4426950408889634f;
    const float t = x * __builtin_fmaf(-kG1 * kL, x * x, -kG0 * kL);
    return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}

__device__ __forceinline__ float gelu_grad(float x) {
    const float t = x * (kG0 + kG1 * (x * x));
    const float s = sigmoid_f(t);
    return s + x * (s * (1.0f - s)) * (kG0 + 3.0f * kG1 * (x * x));
}

__device__ __forceinline__ float in_scale(uint8_t x) { return ((float)x - 128.0f) * (1.0f / 128.0f); }

// ---- training GEMM ---------------------------------------------------------
// C (M x N, row stride ldc) = sum_k A(i, k) B(k, j) with A(i, k) at
// A[i * sai + k * sak] and B(k, j) at B[k * sbk + j * sbj] (any transposes by
// strides).  Sources are FP32, or uint8 patch bytes scaled as the network
// input.  Operands are rounded to BF16; MFMA 32x32x16 accumulates in FP32.
template <bool AU8, bool BU8>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const void *__restrict__ A, long long sai,
                                              long long sak, const void *__restrict__ B, long long sbk,
                                              long long sbj, float *__restrict__ C, long long ldc) {
    __shared__ __align__(16) __bf16 As[64 * 40];   // [m][k], k chunk of 32
    __shared__ __align__(16) __bf16 Bs[64 * 40];   // [n][k]
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wid & 1) * 32, wn = (wid >> 1) * 32;
    floatx16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 32) {
        for (int e = tid; e < 64 * 32; e += 256) {
            const int r = e >> 5, kk = e & 31;
            const int k = k0 + kk;
            const int i = m0 + r, j = n0 + r;
            float va = 0.0f, vb = 0.0f;
            if (i < M && k < K) {
                const long long o = (long long)i * sai + (long long)k * sak;
                va = AU8 ? in_scale(static_cast<const uint8_t *>(A)[o]) : static_cast<const float *>(A)[o];
            }
            if (j < N && k < K) {
                const long long o = (long long)k * sbk + (long long)j * sbj;
                vb = BU8 ? in_scale(static_cast<const uint8_t *>(B)[o]) : static_cast<const float *>(B)[o];
            }
            As[r * 40 + kk] = (__bf16)va;
            Bs[r * 40 + kk] = (__bf16)vb;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 32; ks += 16) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[(wm + (lane & 31)) * 40 + ks + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[(wn + (lane & 31)) * 40 + ks + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) C[(long long)row * ldc + col] = acc[r];
    }
}

hipError_t gemm(int M, int N, int K, const void *A, bool au8, long long sai, long long sak, const void *B, bool bu8,
                long long sbk, long long sbj, float *C, long long ldc, hipStream_t st) {
    const dim3 grid(nblk(N, 64), nblk(M, 64));
    if (au8 && !bu8)
        hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, st, M, N, K, A, sai, sak, B, sbk, sbj, C, ldc);
    else if (!au8 && bu8)
        hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, st, M, N, K, A, sai, sak, B, sbk, sbj, C, ldc);
    else if (!au8 && !bu8)
        hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, st, M, N, K, A, sai, sak, B, sbk, sbj, C,
                           ldc);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Z (n x h) += bias; H = gelu(Z) when H is given
__global__ void k_bias_act(float *__restrict__ Z, const float *__restrict__ bias, float *__restrict__ H,
                           long long n, int h) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * h) return;
    const float z = Z[i] + bias[i % h];
    Z[i] = z;
    if (H) H[i] = gelu_f(z);
}

// dH (n x h) *= gelu'(Zpre)
__global__ void k_gelu_bwd(float *__restrict__ dH, const float *__restrict__ Zpre, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dH[i] *= gelu_grad(Zpre[i]);
}

// out[j] = sum_i X[i][j] (column sums, fixed order: 4 row strides, then the
// four partials in order)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ X, long long n, int h,
                                                float *__restrict__ out) {
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + c;
    float s = 0.0f;
    if (j < h)
        for (long long i = g; i < n; i += 4) s += X[i * h + j];
    part[g][c] = s;
    __syncthreads();
    if (g == 0 && j < h) out[j] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// Triplet margin loss on rows [q; p; n] of Z (3B x d), PyTorch's
// TripletMarginLoss(margin, p=2, eps=1e-6, mean): d(x, y) = ||x - y + eps||.
// Writes per-triplet losses and dZ of the mean loss.
__global__ void k_triplet(const float *__restrict__ Z, int B, int d, float margin, float *__restrict__ dZ,
                          float *__restrict__ losses) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B) return;
    const float eps = 1e-6f;
    const float *zq = Z + (size_t)t * d, *zp = Z + (size_t)(B + t) * d, *zn = Z + (size_t)(2 * B + t) * d;
    float sp = 0.0f, sn = 0.0f;
    for (int k = 0; k < d; ++k) {
        const float a = zq[k] - zp[k] + eps, b = zq[k] - zn[k] + eps;
        sp += a * a;
        sn += b * b;
    }
    const float dp = sqrtf(sp), dn = sqrtf(sn);
    const float l = dp - dn + margin;
    losses[t] = l > 0.0f ? l : 0.0f;
    float *gq = dZ + (size_t)t * d, *gp = dZ + (size_t)(B + t) * d, *gn = dZ + (size_t)(2 * B + t) * d;
    const float inv = 1.0f / (float)B;
    for (int k = 0; k < d; ++k) {
        if (l > 0.0f) {
            const float a = (zq[k] - zp[k] + eps) / dp, b = (zq[k] - zn[k] + eps) / dn;
            gq[k] = (a - b) * inv;
            gp[k] = -a * inv;
            gn[k] = b * inv;
        } else {
            gq[k] = gp[k] = gn[k] = 0.0f;
        }
    }
}

// mean of n values, one block, fixed order
__global__ __launch_bounds__(256) void k_mean(const float *__restrict__ x, int n, float *__restrict__ out) {
    __shared__ float part[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0] / (float)n;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- fused inference --------------------------------------------------------
constexpr int kEncTile = 64;    // patches per tile
constexpr int kHS = 128 + 8;    // hidden tile row stride (bf16)

// Persistent blocks: each wavefront keeps its 32 rows of W1 and W2 in
// registers as MFMA A fragments (W3, used by waves 0-1, sits in LDS); per
// tile of 64 patches: stage the patches as BF16 in LDS (rows padded to whole
// dwords: W1 is stored in the same padded feature order, zero at the pads;
// the rows were loaded during the previous tile), layer 1 as W1 . X^T (wave w
// owns hidden units 32w..32w+31, so each lane holds four consecutive units of
// one patch and writes them as one 8-byte LDS store) + GELU, the MFMAs of the
// second 32 patches scheduled between the GELUs of the first; layer 2 the
// same way over H1 (while the next tile's rows load); layer 3 (d x 64) by
// waves 0-1, written as FP32 rows of dpad floats (zero past d), four floats
// per store.
template <int DK>
__global__ __launch_bounds__(256, 2) void k_encode(PatchGeom g, long long npatch, const __bf16 *__restrict__ W1p,
                                                const float *__restrict__ b1, const __bf16 *__restrict__ W2,
                                                const float *__restrict__ b2, const __bf16 *__restrict__ W3p,
                                                const float *__restrict__ b3p, int d, int dpad,
                                                float *__restrict__ out) {
    constexpr int XS = DK + 8;                          // X tile row stride (bf16)
    constexpr int KS1 = DK / 16;
    extern __shared__ __align__(16) __bf16 esm[];
    __bf16 *X = esm;                                    // kEncTile x XS, later H2 (kEncTile x kHS)
    __bf16 *H1 = esm + kEncTile * XS;                   // kEncTile x kHS
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int lc = lane & 31, lh = lane >> 5;
    const int hcol = 32 * wid + lc;

    __shared__ __align__(16) float sbias[128 + 128 + 32];   // b1, b2, b3 (padded)
    __shared__ __align__(16) __bf16 sw3[32 * kHS];             // W3 (padded to 32 rows)
    bf16x8 w1f[KS1], w2f[8];
#pragma unroll
    for (int ks = 0; ks < KS1; ++ks)
        w1f[ks] = *reinterpret_cast<const bf16x8 *>(W1p + (size_t)hcol * DK + 16 * ks + 8 * lh);
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) w2f[ks] = *reinterpret_cast<const bf16x8 *>(W2 + (size_t)hcol * 128 + 16 * ks + 8 * lh);
    for (int i = tid; i < 32 * 128 / 8; i += 256) {
        const int r = i / 16, c = (i % 16) * 8;
        *reinterpret_cast<bf16x8 *>(sw3 + r * kHS + c) = *reinterpret_cast<const bf16x8 *>(W3p + (size_t)r * 128 + c);
    }
    for (int i = tid; i < 128 + 128 + 32; i += 256) sbias[i] = i < 128 ? b1[i] : (i < 256 ? b2[i - 128] : b3p[i - 256]);
    // rows of this lane's accumulator group g: units 32 * wid + 8 * g + 4 * lh + 0..3
    const int u0 = 32 * wid + 4 * lh;

    const int rw = (g.rowlen + 3) >> 2;
    const int tail = g.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : ~0u;
    const int used = g.rows * rw * 4;                  // padded features staged per patch
    const long long ntiles = (npatch + kEncTile - 1) / kEncTile;

    // a quad of lanes per patch; the rows of the next tile load while the
    // current one runs layers 2 and 3
    using SS = StageShape<DK>;
    const int pl = tid >> 2, q4 = tid & 3;
    uint32_t rows[SS::RL][SS::LW];
    long long o = 0;
    auto fetch = [&](long long tile) {
        const long long pi = tile * kEncTile + pl;
        const bool pv = pi < npatch;
        o = patch_origin(g, pv ? pi : 0);
        fetch_rows(g, o, pv, q4, 4, rows);
    };
    if (blockIdx.x < ntiles) fetch(blockIdx.x);

    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long long p0 = tile * kEncTile;
        // stage
        {
            uint32_t *dst = reinterpret_cast<uint32_t *>(X + (size_t)pl * XS);
            // raw bytes (the packed weights carry the input scaling); integers
            // 0..255 are exact in BF16
            for_row_words(g, o, rows, q4, 4, rw, tmask, [&](int r, int k, uint32_t v) {
                const floatx4 f = {(float)(v & 255u), (float)((v >> 8) & 255u), (float)((v >> 16) & 255u),
                                   (float)(v >> 24)};
                const uint2 h = __builtin_bit_cast(uint2, __builtin_convertvector(f, bf16x4));
                dst[2 * (r * rw + k)] = h.x;
                dst[2 * (r * rw + k) + 1] = h.y;
            });
            // features past the padded rows, up to DK, are zero
            for (int k = used / 2 + q4; k < DK / 2; k += 4) dst[k] = 0u;
        }
        __syncthreads();
        // layer 1: acc[r] = unit 32 * wid + 8 * (r / 4) + 4 * lh + r % 4 of patch 32 * pt + lc;
        // the second half's MFMAs are interleaved with the first half's GELU
        {
            floatx16 acc[2];   // start at the bias
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) {
                const floatx4 bb = *reinterpret_cast<const floatx4 *>(sbias + u0 + 8 * gq);
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[0][4 * gq + i] = acc[1][4 * gq + i] = bb[i];
            }
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                const __bf16 *xb = X + (size_t)(32 * pt + lc) * XS + 8 * lh;
#pragma unroll
                for (int ks = 0; ks < KS1; ++ks)
                    acc[pt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        w1f[ks], *reinterpret_cast<const bf16x8 *>(xb + 16 * ks), acc[pt], 0, 0, 0);
            }
#pragma unroll
            for (int pt = 0; pt < 2; ++pt) {
                __bf16 *hrow = H1 + (size_t)(32 * pt + lc) * kHS;
#pragma unroll
                for (int gq = 0; gq < 4; ++gq) {
                    const floatx4 v = {gelu_fast(acc[pt][4 * gq + 0]), gelu_fast(acc[pt][4 * gq + 1]),
                                       gelu_fast(acc[pt][4 * gq + 2]), gelu_fast(acc[pt][4 * gq + 3])};
                    *reinterpret_cast<bf16x4 *>(hrow + u0 + 8 * gq) = __builtin_convertvector(v, bf16x4);
                }
            }
            // schedule: LDS reads eight ahead; the first half's MFMA chain,
            // then each MFMA of the second half followed by a share of the
            // first half's GELU
            __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);   // LDS reads
#pragma unroll
            for (int i = 0; i < KS1; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
#pragma unroll
            for (int i = 0; i < KS1; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x402, 5, 0);   // VALU incl. transcendental
            }
        }
        __syncthreads();
        if (tile + gridDim.x < ntiles) fetch(tile + gridDim.x);
        // layer 2 (H2 over the X tile)
        floatx16 acc2[2];   // start at the bias
#pragma unroll
        for (int gq = 0; gq < 4; ++gq) {
            const floatx4 bb = *reinterpret_cast<const floatx4 *>(sbias + 128 + u0 + 8 * gq);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc2[0][4 * gq + i] = acc2[1][4 * gq + i] = bb[i];
        }
#pragma unroll
        for (int pt = 0; pt < 2; ++pt) {
            const __bf16 *hb = H1 + (32 * pt + lc) * kHS + 8 * lh;
#pragma unroll
            for (int ks = 0; ks < 8; ++ks)
                acc2[pt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2f[ks], *reinterpret_cast<const bf16x8 *>(hb + 16 * ks),
                                                                   acc2[pt], 0, 0, 0);
        }
#pragma unroll
        for (int pt = 0; pt < 2; ++pt) {
            __bf16 *hrow = X + (size_t)(32 * pt + lc) * kHS;
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) {
                const floatx4 v = {gelu_fast(acc2[pt][4 * gq + 0]), gelu_fast(acc2[pt][4 * gq + 1]),
                                   gelu_fast(acc2[pt][4 * gq + 2]), gelu_fast(acc2[pt][4 * gq + 3])};
                *reinterpret_cast<bf16x4 *>(hrow + u0 + 8 * gq) = __builtin_convertvector(v, bf16x4);
            }
        }
        __syncthreads();
        // layer 3: acc[r] = output 8 * (r / 4) + 4 * lh + r % 4 of patch 32 * wid + lc
        if (wid < 2) {
            floatx16 acc = {};
            const __bf16 *hb = X + (32 * wid + lc) * kHS + 8 * lh;
#pragma unroll
            for (int ks = 0; ks < 8; ++ks)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    *reinterpret_cast<const bf16x8 *>(sw3 + lc * kHS + 16 * ks + 8 * lh),
                    *reinterpret_cast<const bf16x8 *>(hb + 16 * ks), acc, 0, 0, 0);
            const long long pi = p0 + 32 * wid + lc;
            if (pi < npatch) {
#pragma unroll
                for (int gq = 0; gq < 4; ++gq) {
                    const int o = 8 * gq + 4 * lh;
                    if (o < dpad) {
                        floatx4 v;
#pragma unroll
                        for (int i = 0; i < 4; ++i) v[i] = o + i < d ? acc[4 * gq + i] + sbias[256 + o + i] : 0.0f;
                        float *dst = out + pi * dpad + o;
                        if (o + 4 <= dpad && (dpad & 3) == 0) {
                            *reinterpret_cast<floatx4 *>(dst) = v;
                        } else {
                            // a row stride that is not a whole number of
                            // float4s: element stores, none past the row
#pragma unroll
                            for (int i = 0; i < 4; ++i)
                                if (o + i < dpad) dst[i] = v[i];
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
}

// FP32 parameters -> the fused kernel's BF16 layouts
__global__ void k_pack_encoder(const float *__restrict__ params, int p, int h1, int h2, int d, int DK,
                               __bf16 *__restrict__ W1p, float *__restrict__ b1, __bf16 *__restrict__ W2,
                               float *__restrict__ b2, __bf16 *__restrict__ W3p, float *__restrict__ b3p) {
    const int D = p * p * 3, rowlen = 3 * p, rw4 = (rowlen + 3) / 4 * 4;
    const float *pW1 = params, *pb1 = pW1 + (size_t)h1 * D, *pW2 = pb1 + h1, *pb2 = pW2 + (size_t)h2 * h1;
    const float *pW3 = pb2 + h2, *pb3 = pW3 + (size_t)d * h2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (long long)h1 * DK) {
        // the input scaling (x - 128) / 128 moves into the weights: W1 / 128
        // (exact in BF16) applied to the raw bytes, minus sum_k W1 in the bias
        const int j = (int)(i / DK), k = (int)(i % DK);
        const int r = k / rw4, u = k % rw4;
        float w = 0.0f;
        if (r < p && u < rowlen) w = pW1[(size_t)j * D + r * rowlen + u];
        W1p[i] = (__bf16)((float)(__bf16)w * 0.0078125f);
    }
    if (i < (long long)h2 * h1) W2[i] = (__bf16)pW2[i];
    if (i < 32LL * h2) {
        const int j = (int)(i / h2), k = (int)(i % h2);
        W3p[i] = (__bf16)(j < d ? pW3[(size_t)j * h2 + k] : 0.0f);
    }
    if (i < h1) {
        double acc = 0.0;
        for (int f = 0; f < D; ++f) acc += (double)(float)(__bf16)pW1[(size_t)i * D + f];
        b1[i] = (float)((double)pb1[i] - acc);
    }
    if (i < h2) b2[i] = pb2[i];
    if (i < 32) b3p[i] = i < d ? pb3[i] : 0.0f;
}

}  // namespace

long long mlp_param_count(int D, int h1, int h2, int d) {
    return (long long)h1 * D + h1 + (long long)h2 * h1 + h2 + (long long)d * h2 + d;
}

size_t mlp_workspace_floats(long long n, int h1, int h2, int d) {
    // Z1, H1, Z2, H2 (n x h), Z (n x d), dZ (n x d), losses (n / 3 + 1)
    return (size_t)n * (2 * h1 + 2 * h2 + 2 * d) + (size_t)n / 3 + 1;
}

hipError_t mlp_forward(const uint8_t *x, long long n, int D, const float *params, int h1, int h2, int d,
                       float *ws, float *out, hipStream_t st) {
    const float *W1 = params, *b1 = W1 + (size_t)h1 * D, *W2 = b1 + h1, *b2 = W2 + (size_t)h2 * h1;
    const float *W3 = b2 + h2, *b3 = W3 + (size_t)d * h2;
    float *Z1 = ws, *H1 = Z1 + n * h1, *Z2 = H1 + n * h1, *H2 = Z2 + n * h2;
    hipError_t e;
    if ((e = gemm((int)n, h1, D, x, true, D, 1, W1, false, 1, D, Z1, h1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * h1, 256)), dim3(256), 0, st, Z1, b1, H1, n, h1);
    if ((e = gemm((int)n, h2, h1, H1, false, h1, 1, W2, false, 1, h1, Z2, h2, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * h2, 256)), dim3(256), 0, st, Z2, b2, H2, n, h2);
    if ((e = gemm((int)n, d, h2, H2, false, h2, 1, W3, false, 1, h2, out, d, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bias_act, dim3(nblk(n * d, 256)), dim3(256), 0, st, out, b3, (float *)nullptr, n, d);
    return hipGetLastError();
}

hipError_t mlp_triplet_grad(const uint8_t *x, int B, int D, const float *params, int h1, int h2, int d,
                            float margin, float *grads, float *loss, float *ws, hipStream_t st) {
    const long long n = 3LL * B;
    const float *W1 = params, *b1 = W1 + (size_t)h1 * D, *W2 = b1 + h1, *b2 = W2 + (size_t)h2 * h1;
    const float *W3 = b2 + h2;
    (void)b1;
    (void)b2;
    float *gW1 = grads, *gb1 = gW1 + (size_t)h1 * D, *gW2 = gb1 + h1, *gb2 = gW2 + (size_t)h2 * h1;
    float *gW3 = gb2 + h2, *gb3 = gW3 + (size_t)d * h2;
    float *Z1 = ws, *H1 = Z1 + n * h1, *Z2 = H1 + n * h1, *H2 = Z2 + n * h2, *Z = H2 + n * h2, *dZ = Z + n * d;
    float *losses = dZ + n * d;
    hipError_t e;
    if ((e = mlp_forward(x, n, D, params, h1, h2, d, ws, Z, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_triplet, dim3(nblk(B, 256)), dim3(256), 0, st, Z, B, d, margin, dZ, losses);
    if (loss) hipLaunchKernelGGL(k_mean, dim3(1), dim3(256), 0, st, losses, B, loss);
    // layer 3: dW3 = dZ^T H2, db3 = colsum dZ, dH2 = dZ W3 (into H2's buffer after use)
    if ((e = gemm(d, h2, (int)n, dZ, false, 1, d, H2, false, h2, 1, gW3, h2, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(d, 64)), dim3(256), 0, st, dZ, n, d, gb3);
    float *dH2 = H2;
    if ((e = gemm((int)n, h2, d, dZ, false, d, 1, W3, false, h2, 1, dH2, h2, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(n * h2, 256)), dim3(256), 0, st, dH2, Z2, n * h2);
    // layer 2: dW2 = dZ2^T H1, db2, dH1 = dZ2 W2 (into Z2's buffer is not possible: dZ2 lives in H2's)
    if ((e = gemm(h2, h1, (int)n, dH2, false, 1, h2, H1, false, h1, 1, gW2, h1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(h2, 64)), dim3(256), 0, st, dH2, n, h2, gb2);
    float *dH1 = H1;
    if ((e = gemm((int)n, h1, h2, dH2, false, h2, 1, W2, false, h1, 1, dH1, h1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_gelu_bwd, dim3(nblk(n * h1, 256)), dim3(256), 0, st, dH1, Z1, n * h1);
    // layer 1: dW1 = dZ1^T X, db1
    if ((e = gemm(h1, D, (int)n, dH1, false, 1, h1, x, true, D, 1, gW1, D, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_colsum, dim3(nblk(h1, 64)), dim3(256), 0, st, dH1, n, h1, gb1);
    return hipGetLastError();
}

hipError_t mlp_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                    float eps, int step, hipStream_t st) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

// The padded patch length of the k_encode instantiation for patch size p:
// the W1 layout (row stride) the kernel reads, so the packing must use it.
int encoder_dk(int p) {
    const int rw = (3 * p + 3) / 4;
    const int need = p * rw * 4;
    if (need <= 176) return 176;
    if (need <= 256) return 256;
    if (need <= 400) return 400;
    if (need <= 512) return 512;
    return (need + 15) / 16 * 16;   // unsupported (encoder_supported rejects it)
}

bool encoder_supported(int p, int h1, int h2, int d) {
    return p >= 1 && encoder_dk(p) <= 512 && h1 == 128 && h2 == 128 && d >= 1 && d <= 32;
}

hipError_t pack_encoder(const float *params, int p, int h1, int h2, int d, __bf16 *W1p, float *b1, __bf16 *W2,
                        float *b2, __bf16 *W3p, float *b3p, hipStream_t st) {
    const int DK = encoder_dk(p);
    const long long n = (long long)h1 * DK > (long long)h2 * h1 ? (long long)h1 * DK : (long long)h2 * h1;
    hipLaunchKernelGGL(k_pack_encoder, dim3(nblk(n, 256)), dim3(256), 0, st, params, p, h1, h2, d, DK, W1p, b1, W2,
                       b2, W3p, b3p);
    return hipGetLastError();
}

hipError_t encode_patches(const PatchGeom &g, long long npatch, const __bf16 *W1p, const float *b1,
                          const __bf16 *W2, const float *b2, const __bf16 *W3p, const float *b3p, int d, int dpad,
                          float *out, hipStream_t st, int per_cu) {
    if (npatch == 0) return hipSuccess;
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const long long ntiles = (npatch + kEncTile - 1) / kEncTile;
    const long long cap = (long long)(per_cu > 0 ? per_cu : 2) * ncu;
    const unsigned grid = (unsigned)(ntiles < cap ? ntiles : cap);
    auto launch = [&](auto dk) -> hipError_t {
        constexpr int DK = decltype(dk)::value;
        const size_t lds = ((size_t)kEncTile * (DK + 8) + (size_t)kEncTile * kHS) * sizeof(__bf16);
        hipError_t e = hipFuncSetAttribute((const void *)k_encode<DK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_encode<DK>, dim3(grid), dim3(256), lds, st, g, npatch, W1p, b1, W2, b2, W3p, b3p, d,
                           dpad, out);
        return hipGetLastError();
    };
    switch (encoder_dk(g.rows)) {
    case 176: return launch(std::integral_constant<int, 176>{});
    case 256: return launch(std::integral_constant<int, 256>{});
    case 400: return launch(std::integral_constant<int, 400>{});
    case 512: return launch(std::integral_constant<int, 512>{});
    default: return hipErrorInvalidValue;
    }
}

}  // namespace annf
