// Full-patch leaf k-NN with FP8-E4M3 operands (the FP8 precision variant):
// approximate distances on MFMA pick a shortlist per query, which is
// re-ranked with exact integer distances.
//
// Quantization (k_patch_fp8, once per image and call).  The bytes of patch
// i, centred (x - 128), are scaled by s_i = 448 / m_i, m_i = max |x - 128|
// over the patch (s_i = 1 for a flat patch), and rounded to E4M3 (round to
// nearest even).  Each row of 3p bytes is padded to whole dwords with zero
// codes and the patch to DQ bytes, a multiple of 64.  Stored per patch: the
// DQ codes, the inverse scale u_i = m_i / 448 (1 for a flat patch) and the
// norm of the de-scaled codes n_i = (sum_e q_e^2) u_i^2; the sum is exact in
// FP64 (every code is a multiple of 2^-9 below 449) and the product is two
// FP64 roundings of it, then one to FP32.
//
// Approximate squared distance of query a and point b:
//     d~ = n_a + n_b - 2 (q_a . q_b) u_a u_b,
// the dot product of the codes on v_mfma_f32_32x32x64_f8f6f4 with FP32
// accumulation, so d~ is |q_a u_a - q_b u_b|^2 of the de-scaled codes up to
// rounding (a patch against itself scores about 0).
//
// Leaf pass (k_leaf_fp8).  The work split of the BF16 pass: 128 queries of
// one leaf bucket per block, 32 per wavefront with their codes in registers
// (lane l: query l & 31, bytes 32 (l >> 5) .. + 31 of every 64-byte k-step),
// the leaf's points staged 32 at a time in LDS (double buffered).  The key
// of a candidate is the FP32 bits of max(d~, 0) with the low PB bits
// replaced by its position in the leaf (PB bits hold the largest leaf);
// positions follow ascending B index, so keys order like (d~ rounded down,
// B index).  The 32x32 MFMA block leaves lane h of a query the points
// whose position bit 2 is h (16 per tile); each lane keeps the LM / 2
// smallest keys of its half (a v_med3 network; LM = S rounded up to 8, 16
// or 32), and at a leaf's end the two lanes merge their lists (bitonic) and
// the S smallest are written as 64-bit keys (d~ rounded down << 32 | B
// index), empty slots ~0: the S smallest of the leaf whenever neither half
// holds more than LM / 2 of them (always so for its S / 2 smallest).
//
// Shortlist and re-rank (k_fp8_select), 32 lanes per query: lane i holds
// entry i of each of the query's per-leaf lists; pairwise bitonic merges
// across the lanes leave the S smallest of all its scanned leaves, one per
// lane.  Lane i then scores its entry exactly: |a'|^2 + |b'|^2 - 2 a'.b' of
// the centred bytes x' = x - 128 from the uint8 images (v_dot4_i32_i8; the
// query's rows aligned once into LDS, the candidate's row windows loaded
// together; the centred norms precomputed per patch).  The 32 exact keys
// (d2 << 32 | B index) are sorted across the lanes and merged with any exact
// keys given for the query (the natural pass's list in the propagated pass);
// the K smallest are the answer, the order of the exact field.
#include "annf_internal.h"
#include "annf_patch.h"
#include "annf_lists.h"

#include <type_traits>

namespace annf {

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef int intx8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));

constexpr int kQ = 128;          // queries per block (4 wavefronts x 32)
constexpr int kTP = 32;          // points per staged tile
constexpr int kThreads = 256;
constexpr int kStageSets = kThreads / kTP;   // staging threads per point
constexpr int kMaxDQ = 448;      // p <= 12: 12 rows of 9 dwords, rounded up to 64
constexpr uint64_t kNoKey = ~0ull;

// (x - 128) * scale rounded to E4M3 (round to nearest even), byte by byte;
// bytes outside `mask` become code 0.
__device__ __forceinline__ uint32_t fp8_word(uint32_t w, float scale, uint32_t mask) {
    const float a0 = ((float)(w & 255u) - 128.0f) * scale;
    const float a1 = ((float)((w >> 8) & 255u) - 128.0f) * scale;
    const float a2 = ((float)((w >> 16) & 255u) - 128.0f) * scale;
    const float a3 = ((float)(w >> 24) - 128.0f) * scale;
    int packed = __builtin_amdgcn_cvt_pk_fp8_f32(a0, a1, 0, false);
    packed = __builtin_amdgcn_cvt_pk_fp8_f32(a2, a3, packed, true);
    return (uint32_t)packed & mask;
}

__device__ __forceinline__ int quad_max_i(int v) {
    v = max(v, __shfl_xor(v, 1, 64));
    v = max(v, __shfl_xor(v, 2, 64));
    return v;
}

__device__ __forceinline__ double quad_sum_d(double v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    return v;
}

// ---------------------------------------------------------------------------
// quantization
// ---------------------------------------------------------------------------

constexpr int kQuantPatches = 64;   // patches per block (a quad of lanes each)

// A quad of lanes per patch (rows q4, q4 + 4, ...).  The codes go through
// LDS (row stride dq / 4 + 4 words) so that the block writes its patches'
// rows with coalesced 16-byte stores.
template <int RL, int LW>
__global__ __launch_bounds__(256) void k_patch_fp8(PatchGeom g, long long n, int dq, uint8_t *__restrict__ q,
                                                   float *__restrict__ nrm, float *__restrict__ inv) {
    extern __shared__ __align__(16) uint32_t sq[];
    const int tid = threadIdx.x, pl = tid >> 2, q4 = tid & 3;
    const int dw = dq / 4, stride = dw + 4;
    const long long p0 = (long long)blockIdx.x * kQuantPatches;
    const long long i = p0 + pl;
    const bool valid = i < n;
    const long long o = patch_origin(g, valid ? i : 0);
    uint32_t w[RL][LW];
    fetch_rows(g, o, valid, q4, 4, w);
    const int rw = (g.rowlen + 3) >> 2;
    const int tail = g.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : ~0u;
    int mx = 0;
    for_row_words(g, o, w, q4, 4, rw, tmask, [&](int, int k, uint32_t v) {
        const uint32_t m = k == rw - 1 ? tmask : ~0u;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if ((m >> (8 * b)) & 255u) mx = max(mx, abs((int)((v >> (8 * b)) & 255u) - 128));
    });
    mx = quad_max_i(mx);
    const float sc = mx ? 448.0f / (float)mx : 1.0f;
    const float iv = mx ? (float)mx / 448.0f : 1.0f;
    uint32_t *row = sq + pl * stride;
    double ss = 0.0;
    for_row_words(g, o, w, q4, 4, rw, tmask, [&](int r, int k, uint32_t v) {
        const uint32_t c = fp8_word(v, sc, k == rw - 1 ? tmask : ~0u);
        row[r * rw + k] = c;
        const double f0 = __builtin_amdgcn_cvt_f32_fp8((int)c, 0), f1 = __builtin_amdgcn_cvt_f32_fp8((int)c, 1);
        const double f2 = __builtin_amdgcn_cvt_f32_fp8((int)c, 2), f3 = __builtin_amdgcn_cvt_f32_fp8((int)c, 3);
        ss += f0 * f0;
        ss += f1 * f1;
        ss += f2 * f2;
        ss += f3 * f3;
    });
    for (int k = g.rows * rw + q4; k < dw; k += 4) row[k] = 0u;
    ss = quad_sum_d(ss);
    if (valid && q4 == 0) {
        nrm[i] = (float)(ss * ((double)iv * (double)iv));
        inv[i] = iv;
    }
    __syncthreads();
    const int np = (int)min((long long)kQuantPatches, n - p0);
    u32x4 *dst = reinterpret_cast<u32x4 *>(q + p0 * dq);
    for (int e = tid; e < np * (dw / 4); e += 256) {
        const int r = e / (dw / 4), c = e - r * (dw / 4);
        dst[e] = *reinterpret_cast<const u32x4 *>(sq + r * stride + 4 * c);
    }
}

// ---------------------------------------------------------------------------
// leaf pass
// ---------------------------------------------------------------------------

struct Fp8Patches {
    const uint8_t *q;     // n x dq codes
    const float *nrm;     // n norms of the de-scaled codes
    const float *inv;     // n inverse scales
};

template <int KS, int LM>
__global__ __launch_bounds__(kThreads, 2) void k_leaf_fp8(TreeView t, Fp8Patches A, Fp8Patches B,
                                                        const int *__restrict__ pts_order,
                                                        const int *__restrict__ order,
                                                        const int *__restrict__ chunk_off,
                                                        const int4 *__restrict__ blk_info, int nprobe, int SP,
                                                        int slot_div, int a0, int pb, uint64_t *__restrict__ keys) {
    constexpr int CH = 4 * KS;                                  // 16-byte chunks per point
    constexpr int DQ = 64 * KS;
    constexpr int CPT = (CH + kStageSets - 1) / kStageSets;     // chunks per staging thread
    __shared__ __align__(16) uint8_t bt[2][CH * kTP * 16];     // [chunk][point][16 bytes]
    __shared__ __align__(16) float pnb[2][kTP], pm2[2][kTP];
    __shared__ __align__(16) uint32_t ppos[2][kTP];

    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int h = lane >> 5, qc = lane & 31;
    const int nleaves = 1 << t.depth;
    const int b = blockIdx.x;
    if (b >= chunk_off[nleaves]) return;
    const int4 bi0 = blk_info[2 * b];
    const int leaf = bi0.x, qb = bi0.y, nqb = bi0.z;
    if (nqb <= 0) return;

    const int qi = 32 * wid + qc;
    const bool qv = qi < nqb;
    const bool wave_live = 32 * wid < nqb;
    const int slot = qv ? order[qb + qi] : 0;
    const uint32_t pmask = (1u << pb) - 1u;

    // the probed leaves' outputs: an empty leaf's lists are written here
    // (the tile walk below skips it)
    for (int pr = 0; pr < nprobe; ++pr) {
        const int pl = leaf ^ pr;
        if (t.leaves[2 * pl + 1] == t.leaves[2 * pl] && qv && h == 0)
            for (int e = 0; e < SP; ++e) keys[((size_t)slot * nprobe + pr) * SP + e] = kNoKey;
    }

    // the query operand: bytes 64 s + 32 h .. + 31 of its codes
    intx8 af[KS];
    float na = 0.0f, sa = 0.0f;
    {
        const int ap = a0 + slot / slot_div;
        if (qv) {
            na = A.nrm[ap];
            sa = A.inv[ap];
        }
        const u32x4 *src = reinterpret_cast<const u32x4 *>(A.q + (size_t)ap * DQ + 32 * h);
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const u32x4 lo = src[4 * s], hi = src[4 * s + 1];
            af[s] = intx8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1], (int)hi[2],
                          (int)hi[3]};
        }
    }

    // tile walk over the probed leaves (empty ones skipped)
    int pr = 0, from = t.leaves[2 * leaf], cnt = t.leaves[2 * leaf + 1] - from, t0 = 0;
    auto seek = [&]() -> bool {
        while (t0 >= cnt) {
            if (++pr >= nprobe) return false;
            const int pl = leaf ^ pr;
            from = t.leaves[2 * pl];
            cnt = t.leaves[2 * pl + 1] - from;
            t0 = 0;
        }
        return true;
    };
    // staging slot of this thread: point sp, chunks sc, sc + kStageSets, ...
    const int sp = tid / kStageSets, sc = tid % kStageSets;
    u32x4 raw[CPT];
    float rn = 0.0f, ri = 0.0f;
    uint32_t rpos = 0xFFFFFFFFu;
    int np = 0;
    auto fetch = [&]() {
        np = min(kTP, cnt - t0);
        const bool v = sp < np;
        const int j = v ? pts_order[from + t0 + sp] : 0;
        const u32x4 *src = reinterpret_cast<const u32x4 *>(B.q + (size_t)j * DQ);
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int c = sc + kStageSets * i;
            raw[i] = (v && c < CH) ? src[c] : u32x4{0u, 0u, 0u, 0u};
        }
        if (sc == 0) {
            rn = v ? B.nrm[j] : 0.0f;
            ri = v ? B.inv[j] : 0.0f;
            rpos = v ? (uint32_t)(t0 + sp) : 0xFFFFFFFFu;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int c = sc + kStageSets * i;
            if (c < CH) *reinterpret_cast<u32x4 *>(bt[buf] + (c * kTP + sp) * 16) = raw[i];
        }
        if (sc == 0) {
            pnb[buf][sp] = rn;
            pm2[buf][sp] = -2.0f * ri;
            ppos[buf][sp] = rpos;
        }
    };

    // each lane keeps the LH smallest keys of its half of the points (the
    // tile rows with position bit 2 equal to h); the two halves' lists make
    // the (query, leaf) list of LM
    constexpr int LH = LM / 2;
    uint32_t l32[LH];
#pragma unroll
    for (int s = 0; s < LH; ++s) l32[s] = 0xFFFFFFFFu;

    bool more = seek();
    if (more) fetch();
    int cur = 0;
    if (more) store(0);
    __syncthreads();
    while (more) {
        // this tile: probe pr, points t0 .. t0 + np - 1 of its leaf, in bt[cur]
        const int tpr = pr, tnp = np, tt0 = t0, lfrom = from;
        const bool leaf_end = tt0 + tnp >= cnt;
        t0 += kTP;
        more = seek();
        if (more) fetch();   // in flight during the compute
        if (wave_live) {
            floatx16 acc = {}, acc1 = {};
            const uint8_t *br = bt[cur] + ((2 * h) * kTP + qc) * 16;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const u32x4 lo = *reinterpret_cast<const u32x4 *>(br + (4 * s) * kTP * 16);
                const u32x4 hi = *reinterpret_cast<const u32x4 *>(br + (4 * s + 1) * kTP * 16);
                const intx8 bv = intx8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3], (int)hi[0], (int)hi[1],
                                       (int)hi[2], (int)hi[3]};
                if (s & 1)
                    acc1 = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(bv, af[s], acc1, 0, 0, 0, 0, 0, 0);
                else
                    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(bv, af[s], acc, 0, 0, 0, 0, 0, 0);
            }
            acc += acc1;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // rows 8 j + 4 h + 0..3 of the tile
                const float4 nb4 = *reinterpret_cast<const float4 *>(&pnb[cur][8 * j + 4 * h]);
                const float4 m24 = *reinterpret_cast<const float4 *>(&pm2[cur][8 * j + 4 * h]);
                const u32x4 ps4 = *reinterpret_cast<const u32x4 *>(&ppos[cur][8 * j + 4 * h]);
                const float nbv[4] = {nb4.x, nb4.y, nb4.z, nb4.w};
                const float m2v[4] = {m24.x, m24.y, m24.z, m24.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float d = __builtin_fmaf(acc[4 * j + i], sa * m2v[i], na + nbv[i]);
                    const uint32_t bits = __float_as_uint(fmaxf(d, 0.0f));
                    list_insert<LH>(l32, (bits & ~pmask) | ps4[i]);
                }
            }
            if (leaf_end) {
                // the two lanes' lists sorted together (this lane's list
                // followed by the other's reversed is bitonic; its merging
                // network sorts it), and the S smallest written
                uint32_t c[LM];
#pragma unroll
                for (int s = 0; s < LH; ++s) {
                    c[s] = l32[s];
                    c[LM - 1 - s] = (uint32_t)__shfl_xor((int)l32[s], 32, 64);
                }
#pragma unroll
                for (int k = LM / 2; k >= 1; k >>= 1) {
#pragma unroll
                    for (int s = 0; s < LM; ++s) {
                        if ((s & k) == 0) {
                            const uint32_t x = c[s], y = c[s + k];
                            c[s] = min(x, y);
                            c[s + k] = max(x, y);
                        }
                    }
                }
                if (qv) {
                    // this lane's half of the list as (d~ bits << 32 | B index)
                    uint64_t *dst = keys + ((size_t)slot * nprobe + tpr) * SP;
                    const int *lp = pts_order + lfrom;
#pragma unroll
                    for (int e = 0; e < LM; e += 2) {
                        if (e < SP && ((e < SP / 2) != (h == 1))) {
                            const uint64_t k0 = c[e] == 0xFFFFFFFFu
                                                    ? kNoKey
                                                    : ((uint64_t)(c[e] & ~pmask) << 32) | (uint32_t)lp[c[e] & pmask];
                            const uint64_t k1 =
                                c[e + 1] == 0xFFFFFFFFu
                                    ? kNoKey
                                    : ((uint64_t)(c[e + 1] & ~pmask) << 32) | (uint32_t)lp[c[e + 1] & pmask];
                            *reinterpret_cast<ulonglong2 *>(dst + e) = make_ulonglong2(k0, k1);
                        }
                    }
                }
#pragma unroll
                for (int s = 0; s < LH; ++s) l32[s] = 0xFFFFFFFFu;
            }
        }
        if (more) store(cur ^ 1);
        __syncthreads();   // the next tile is staged; this one is consumed
        cur ^= 1;
    }
}

// ---------------------------------------------------------------------------
// shortlist and exact re-rank
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Two lists sorted ascending across the 32 lanes of a half-wave (lane i
// holds entry i): the 32 smallest of both, sorted the same way.  The first
// list against the second reversed is bitonic after a lane-wise minimum;
// five compare-exchange steps across lanes sort it.
__device__ __forceinline__ uint64_t merge32(uint64_t a, uint64_t b, int lane) {
    const int i = lane & 31;
    uint64_t c = shfl64(b, (lane & 32) | (31 - i));
    c = a < c ? a : c;
#pragma unroll
    for (int k = 16; k >= 1; k >>= 1) {
        const uint64_t x = shfl_xor64(c, k);
        c = (i & k) ? (x > c ? x : c) : (x < c ? x : c);
    }
    return c;
}

// Row window of LW dwords from the dword below byte offset `off` (whole
// 16-byte loads, or bytes near the end of the buffer).
template <int LW>
__device__ __forceinline__ void row_window(const PatchGeom &g, long long off, uint32_t (&w)[LW]) {
    const long long al = off & ~3LL;
    if (al + 4LL * LW <= g.nbytes) {
        const u32x4a *s4 = reinterpret_cast<const u32x4a *>(g.base + al);
#pragma unroll
        for (int k = 0; k < LW / 4; ++k) {
            const u32x4 v = s4[k];
            w[4 * k] = v[0];
            w[4 * k + 1] = v[1];
            w[4 * k + 2] = v[2];
            w[4 * k + 3] = v[3];
        }
    } else {
#pragma unroll
        for (int k = 0; k < LW; ++k) {
            uint32_t v = 0;
            for (int b = 0; b < 4; ++b) {
                const long long o = al + 4 * k + b;
                if (o < g.nbytes) v |= (uint32_t)g.base[o] << (8 * b);
            }
            w[k] = v;
        }
    }
}

// dword k of a patch row at byte offset `off`, aligned (bytes past the end
// of the buffer read as zero)
__device__ __forceinline__ uint32_t row_word(const PatchGeom &g, long long off, int k) {
    const long long al = (off & ~3LL) + 4LL * k;
    const int sh = (int)(off & 3);
    if (al + 8 <= g.nbytes) {
        const uint32_t *p = reinterpret_cast<const uint32_t *>(g.base + al);
        return __builtin_amdgcn_alignbyte(p[1], p[0], sh);
    }
    uint32_t v = 0;
    for (int b = 0; b < 4; ++b) {
        const long long o = al + sh + b;
        if (o < g.nbytes) v |= (uint32_t)g.base[o] << (8 * b);
    }
    return v;
}

constexpr int kSelQ = 8;                  // queries per block (32 lanes each)
constexpr int kMaxRowWords = 12 * 9;      // p <= 12: 12 rows of at most 9 dwords

// Queries q0 .. q0 + nq - 1, 32 lanes each.  Source r (r < nsrc) of query q
// is the list of slot (q - q0) * nsrc + r, its leaf natleaf[q] ^ r
// (pairleaf null) or pairleaf[slot] (-1: not scanned).  RB rows of the
// candidate's patch are loaded together (LW-dword windows).
template <int NS, int LW, int RB>
__global__ __launch_bounds__(256) void k_fp8_select(const uint64_t *__restrict__ keys, int SP, int S, int q0, int nq,
                                                    const int *__restrict__ qorder, int nsrc,
                                                    const int *__restrict__ pairleaf, PatchGeom ga,
                                                    PatchGeom gb, const uint32_t *__restrict__ norm_a,
                                                    const uint32_t *__restrict__ norm_b,
                                                    const uint64_t *__restrict__ base, int K,
                                                    uint64_t *__restrict__ keys_out, int *__restrict__ knn,
                                                    int *__restrict__ nn, float *__restrict__ dist) {
    __shared__ uint32_t sa[kSelQ][kMaxRowWords];
    const int tid = threadIdx.x, lane = tid & 63, i = lane & 31, g = tid >> 5;
    // consecutive groups of queries run on one XCD (blocks are dealt to the
    // XCDs round-robin; the grid is a multiple of kXcds), so the candidates
    // that neighbouring queries share stay in that XCD's L2
    const int per = (int)gridDim.x / kXcds;
    const int lb = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    if ((long long)lb * kSelQ >= nq) return;
    const int pos = lb * kSelQ + g;
    const bool live = pos < nq;         // uniform over the 32 lanes of a query
    const int ql = live ? (qorder ? qorder[pos] - q0 : pos) : 0;
    const int q = q0 + ql;
    const int rw = (ga.rowlen + 3) >> 2;
    const int nw = ga.rows * rw;
    const int tail = ga.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : ~0u;

    // the query's rows, aligned, into LDS
    const long long oa = patch_origin(ga, q);
    for (int e = i; e < nw; e += 32) {
        const int r = e / rw, k = e - r * rw;
        sa[g][e] = (row_word(ga, oa + r * ga.rstride, k) ^ 0x80808080u) & (k == rw - 1 ? tmask : ~0u);
    }

    // the shortlist: entry i of every source list, merged across the lanes
    uint64_t m = kNoKey;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        uint64_t v = kNoKey;
        if (live && r < nsrc && i < S) {
            const size_t slot = (size_t)ql * nsrc + r;
            const uint64_t k = keys[slot * SP + i];
            if (!pairleaf || pairleaf[slot] >= 0) v = k;   // a pair slot of leaf -1 was never scanned
        }
        m = r == 0 ? v : merge32(m, v, lane);
    }
    if (i >= S) m = kNoKey;
    __syncthreads();

    // exact distance of entry i: centred bytes (x ^ 0x80 is x - 128 as a
    // signed byte) on v_dot4_i32_i8, with the patches' centred norms
    uint64_t key = kNoKey;
    if (m != kNoKey) {
        const int j = (int)(uint32_t)m;
        const long long ob = patch_origin(gb, j);
        int sab = 0;
        for (int r0 = 0; r0 < ga.rows; r0 += RB) {
            uint32_t wb[RB][LW];
#pragma unroll
            for (int rr = 0; rr < RB; ++rr) {
                if (r0 + rr < ga.rows) {
                    row_window<LW>(gb, ob + (r0 + rr) * gb.rstride, wb[rr]);
                } else {
#pragma unroll
                    for (int k = 0; k < LW; ++k) wb[rr][k] = 0u;
                }
            }
#pragma unroll
            for (int rr = 0; rr < RB; ++rr) {
                const int r = r0 + rr;
                const int sh = (int)((ob + r * gb.rstride) & 3);
#pragma unroll
                for (int k = 0; k < LW - 1; ++k) {
                    if (r < ga.rows && k < rw) {
                        const int x = (int)sa[g][r * rw + k];
                        const int y = (int)((__builtin_amdgcn_alignbyte(wb[rr][k + 1], wb[rr][k], sh) ^ 0x80808080u) &
                                            (k == rw - 1 ? tmask : ~0u));
                        sab = __builtin_amdgcn_sdot4(x, y, sab, false);
                    }
                }
            }
        }
        key = ((uint64_t)(norm_a[q] + norm_b[j] - 2u * (uint32_t)sab) << 32) | (uint32_t)j;
    }
    // the 32 entries sorted across the lanes (bitonic), then merged with the
    // query's given keys (lanes 0 .. K - 1): lane r holds the r-th smallest
    {
#pragma unroll
        for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
            for (int jj = k >> 1; jj >= 1; jj >>= 1) {
                const uint64_t x = shfl_xor64(key, jj);
                const bool keep_min = ((i & jj) == 0) == ((i & k) == 0);
                key = keep_min ? (x < key ? x : key) : (x > key ? x : key);
            }
        }
    }
    if (base) {
        const uint64_t bk = (live && i < K) ? base[(size_t)q * K + i] : kNoKey;
        key = merge32(key, bk, lane);
    }
    if (live && i < K) {
        const int jj = key == kNoKey ? -1 : (int)(uint32_t)key;
        if (keys_out) keys_out[(size_t)q * K + i] = key;
        if (knn) knn[(size_t)q * K + i] = jj;
        if (i == 0) {
            if (nn) nn[q] = jj;
            if (dist) dist[q] = key == kNoKey ? kFltMax : (float)__builtin_sqrt((double)(uint32_t)(key >> 32));
        }
    }
}

// The two-lanes-per-entry form of k_fp8_select for rows of at most 28
// bytes (p <= 9): a wavefront per query, lanes 2c and 2c + 1 hold entry c.
// Each reads one 16-byte half of the candidate's 32-byte row window (the
// dword past the first half comes from the partner lane), so a load
// instruction touches one cache line per entry instead of one per lane and
// half; the two partial dot products are summed across the pair.
__device__ __forceinline__ uint64_t merge32x2(uint64_t a, uint64_t b, int lane) {
    const int c = lane >> 1;
    uint64_t o = shfl64(b, ((31 - c) << 1) | (lane & 1));
    o = a < o ? a : o;
#pragma unroll
    for (int k = 16; k >= 1; k >>= 1) {
        const uint64_t x = shfl_xor64(o, k << 1);
        o = (c & k) ? (x > o ? x : o) : (x < o ? x : o);
    }
    return o;
}

template <int NS>
__global__ __launch_bounds__(256) void k_fp8_select2(const uint64_t *__restrict__ keys, int SP, int S, int q0, int nq,
                                                     const int *__restrict__ qorder, int nsrc,
                                                     const int *__restrict__ pairleaf, PatchGeom ga, PatchGeom gb,
                                                     const uint32_t *__restrict__ norm_a,
                                                     const uint32_t *__restrict__ norm_b,
                                                     const uint64_t *__restrict__ base, int K,
                                                     uint64_t *__restrict__ keys_out, int *__restrict__ knn,
                                                     int *__restrict__ nn, float *__restrict__ dist) {
    constexpr int kQB = 4;                          // queries per block (a wavefront each)
    __shared__ uint32_t sa[kQB][kMaxRowWords];
    const int tid = threadIdx.x, lane = tid & 63, g = tid >> 6;
    const int c = lane >> 1, part = lane & 1;
    const int per = (int)gridDim.x / kXcds;
    const int lb = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    if ((long long)lb * kQB >= nq) return;
    const int pos = lb * kQB + g;
    const bool live = pos < nq;          // uniform over the wavefront
    const int ql = live ? (qorder ? qorder[pos] - q0 : pos) : 0;
    const int q = q0 + ql;
    const int rw = (ga.rowlen + 3) >> 2;
    const int nw = ga.rows * rw;
    const int tail = ga.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : ~0u;

    const long long oa = patch_origin(ga, q);
    for (int e = lane; e < nw; e += 64) {
        const int r = e / rw, k = e - r * rw;
        sa[g][e] = (row_word(ga, oa + r * ga.rstride, k) ^ 0x80808080u) & (k == rw - 1 ? tmask : ~0u);
    }

    // the shortlist: entry c of every source list (both lanes of a pair)
    uint64_t m = kNoKey;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        uint64_t v = kNoKey;
        if (live && r < nsrc && c < S) {
            const size_t slot = (size_t)ql * nsrc + r;
            const uint64_t kk = keys[slot * SP + c];
            if (!pairleaf || pairleaf[slot] >= 0) v = kk;
        }
        m = r == 0 ? v : merge32x2(m, v, lane);
    }
    if (c >= S) m = kNoKey;
    __syncthreads();

    uint64_t key = kNoKey;
    {
        const bool have = m != kNoKey;
        const int j = have ? (int)(uint32_t)m : 0;
        const long long ob = patch_origin(gb, j);
        int sab = 0;
        for (int r = 0; r < ga.rows; ++r) {
            const long long off = ob + r * gb.rstride;
            const long long al = (off & ~3LL) + 16 * part;
            uint32_t w[4];
            if (al + 16 <= gb.nbytes) {
                const u32x4 v = *reinterpret_cast<const u32x4a *>(gb.base + al);
                w[0] = v[0];
                w[1] = v[1];
                w[2] = v[2];
                w[3] = v[3];
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint32_t v = 0;
                    for (int b = 0; b < 4; ++b) {
                        const long long o = al + 4 * k + b;
                        if (o < gb.nbytes) v |= (uint32_t)gb.base[o] << (8 * b);
                    }
                    w[k] = v;
                }
            }
            const uint32_t w4 = (uint32_t)__shfl_xor((int)w[0], 1, 64);   // the partner's first dword
            const int sh = (int)(off & 3);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int kw = 4 * part + k;                // word of the row
                if (have && kw < rw) {
                    const uint32_t hi = k < 3 ? w[k + 1] : w4;
                    const int y = (int)((__builtin_amdgcn_alignbyte(hi, w[k], sh) ^ 0x80808080u) &
                                        (kw == rw - 1 ? tmask : ~0u));
                    sab = __builtin_amdgcn_sdot4((int)sa[g][r * rw + kw], y, sab, false);
                }
            }
        }
        sab += __shfl_xor(sab, 1, 64);
        if (have) key = ((uint64_t)(norm_a[q] + norm_b[j] - 2u * (uint32_t)sab) << 32) | (uint32_t)j;
    }
    // sorted across the 32 entries, merged with the query's given keys
#pragma unroll
    for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
        for (int jj = k >> 1; jj >= 1; jj >>= 1) {
            const uint64_t x = shfl_xor64(key, jj << 1);
            const bool keep_min = ((c & jj) == 0) == ((c & k) == 0);
            key = keep_min ? (x < key ? x : key) : (x > key ? x : key);
        }
    }
    if (base) {
        const uint64_t bk = (live && c < K) ? base[(size_t)q * K + c] : kNoKey;
        key = merge32x2(key, bk, lane);
    }
    if (live && part == 0 && c < K) {
        const int jj = key == kNoKey ? -1 : (int)(uint32_t)key;
        if (keys_out) keys_out[(size_t)q * K + c] = key;
        if (knn) knn[(size_t)q * K + c] = jj;
        if (c == 0) {
            if (nn) nn[q] = jj;
            if (dist) dist[q] = key == kNoKey ? kFltMax : (float)__builtin_sqrt((double)(uint32_t)(key >> 32));
        }
    }
}

template <typename Fn>
hipError_t with_ks_lm(int ks, int lm_need, Fn &&fn) {
    auto pick_lm = [&](auto k) -> hipError_t {
        if (lm_need <= 8) return fn(k, std::integral_constant<int, 8>{});
        if (lm_need <= 16) return fn(k, std::integral_constant<int, 16>{});
        return fn(k, std::integral_constant<int, 32>{});
    };
    switch (ks) {
    case 1: return pick_lm(std::integral_constant<int, 1>{});
    case 2: return pick_lm(std::integral_constant<int, 2>{});
    case 3: return pick_lm(std::integral_constant<int, 3>{});
    case 4: return pick_lm(std::integral_constant<int, 4>{});
    case 5: return pick_lm(std::integral_constant<int, 5>{});
    case 6: return pick_lm(std::integral_constant<int, 6>{});
    case 7: return pick_lm(std::integral_constant<int, 7>{});
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

int leaffp8_dq(int p) {
    const int rw = (3 * p + 3) / 4;
    return (p * rw * 4 + 63) / 64 * 64;
}

bool leaffp8_supported(int p) { return p >= 1 && leaffp8_dq(p) <= kMaxDQ; }

int leaffp8_stride(int S) { return (S + 7) / 8 * 8; }

hipError_t patch_fp8(const uint8_t *img, int h, int w, int p, uint8_t *q, float *nrm, float *inv, hipStream_t st) {
    if (!leaffp8_supported(p) || h < p || w < p) return hipErrorInvalidValue;
    const long long n = (long long)(h - p + 1) * (w - p + 1);
    if (n == 0) return hipSuccess;
    const int dq = leaffp8_dq(p);
    PatchGeom g{};
    g.base = img;
    g.image = 1;
    g.ncols = w - p + 1;
    g.width = w;
    g.rows = p;
    g.rowlen = 3 * p;
    g.rstride = 3LL * w;
    g.nbytes = 3LL * h * w;
    const size_t lds = (size_t)kQuantPatches * (dq / 4 + 4) * sizeof(uint32_t);
    const unsigned nblk = (unsigned)((n + kQuantPatches - 1) / kQuantPatches);
    // rows per lane (a quad per patch) and window dwords (rw + 1)
    if (p <= 8) {
        hipLaunchKernelGGL((k_patch_fp8<2, 7>), dim3(nblk), dim3(256), lds, st, g, n, dq, q, nrm, inv);
    } else {
        hipLaunchKernelGGL((k_patch_fp8<3, 10>), dim3(nblk), dim3(256), lds, st, g, n, dq, q, nrm, inv);
    }
    return hipGetLastError();
}

hipError_t leaf_fp8(const TreeView &t, const uint8_t *qa, const float *na, const float *ia, const uint8_t *qb,
                    const float *nb, const float *ib, int p, const int *pts_order, const int *order,
                    const int *offset, const int *chunk_off, void *scratch, int nq, int nprobe, int S, int slot_div,
                    int a0, int pb, uint64_t *keys, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    if (!leaffp8_supported(p)) return hipErrorInvalidValue;
    const int nleaves = 1 << t.depth;
    if (nprobe < 1 || nprobe > nleaves || S < 1 || S > 32 || slot_div < 1 || pb < 1 || pb > 24)
        return hipErrorInvalidValue;
    const int ks = leaffp8_dq(p) / 64;
    const unsigned nblk = (unsigned)(nleaves + (nq + kQ - 1) / kQ);
    int4 *blk_info = static_cast<int4 *>(scratch);
    hipError_t e = leaf_blocks_q(chunk_off, offset, t.leaves, nleaves, (int)nblk, blk_info, st);
    if (e != hipSuccess) return e;
    const Fp8Patches A{qa, na, ia}, B{qb, nb, ib};
    const int SP = leaffp8_stride(S);
    return with_ks_lm(ks, SP, [&](auto k, auto lm) -> hipError_t {
        constexpr int KS = decltype(k)::value;
        constexpr int LM = decltype(lm)::value;
        hipLaunchKernelGGL((k_leaf_fp8<KS, LM>), dim3(nblk), dim3(kThreads), 0, st, t, A, B, pts_order, order,
                           chunk_off, blk_info, nprobe, SP, slot_div, a0, pb, keys);
        return hipGetLastError();
    });
}

hipError_t fp8_select(const uint64_t *keys, int S, int q0, int nq, const int *qorder, int nsrc, const int *pairleaf,
                      const PatchGeom &ga, const PatchGeom &gb, const uint32_t *norm_a, const uint32_t *norm_b,
                      const uint64_t *base, int K, uint64_t *keys_out, int *knn, int *nn, float *dist,
                      hipStream_t st) {
    if (nq == 0) return hipSuccess;
    if (nsrc < 1 || nsrc > 32 || S < 1 || S > 32 || K < 1 || K > 32 || ga.rows != gb.rows || ga.rowlen != gb.rowlen ||
        ga.rows * ((ga.rowlen + 3) / 4) > kMaxRowWords)
        return hipErrorInvalidValue;
    const int SP = leaffp8_stride(S);
    const int rw = (ga.rowlen + 3) / 4;
    const long long nblk = ((long long)nq + kSelQ - 1) / kSelQ;
    const dim3 g((unsigned)((nblk + kXcds - 1) / kXcds * kXcds)), b(256);
    // rows of at most 28 bytes: two lanes per entry, a wavefront per query
    const long long nblk2 = ((long long)nq + 3) / 4;
    const dim3 g2((unsigned)((nblk2 + kXcds - 1) / kXcds * kXcds));
    auto launch = [&](auto ns) -> hipError_t {
        constexpr int NS = decltype(ns)::value;
        if (rw <= 7)
            hipLaunchKernelGGL((k_fp8_select2<NS>), g2, b, 0, st, keys, SP, S, q0, nq, qorder, nsrc, pairleaf, ga, gb,
                               norm_a, norm_b, base, K, keys_out, knn, nn, dist);
        else
            hipLaunchKernelGGL((k_fp8_select<NS, 12, 6>), g, b, 0, st, keys, SP, S, q0, nq, qorder, nsrc,
                               pairleaf, ga, gb, norm_a, norm_b, base, K, keys_out, knn, nn, dist);
        return hipGetLastError();
    };
    if (nsrc <= 1) return launch(std::integral_constant<int, 1>{});
    if (nsrc <= 2) return launch(std::integral_constant<int, 2>{});
    if (nsrc <= 4) return launch(std::integral_constant<int, 4>{});
    if (nsrc <= 8) return launch(std::integral_constant<int, 8>{});
    if (nsrc <= 16) return launch(std::integral_constant<int, 16>{});
    return launch(std::integral_constant<int, 32>{});
}

}  // namespace annf
