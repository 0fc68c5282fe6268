// Full-patch leaf k-NN on MFMA (the precision variants and the learned-
// encoder path).  For the queries bucketed by natural leaf, the distance to
// every point of the probed leaves is |a|^2 + |b|^2 - 2 a.b over the full
// p*p*3 uint8 patch, the dot products on MFMA with FP32 accumulation:
//
//  * BF16: bytes 0..255 are exact in BF16 and so is every product; the
//    k-steps alternate between two FP32 accumulators, each summing at most
//    16 steps of 16 products below 65,025 (D <= 512), so no partial sum
//    reaches 2^24, both halves are exact integers and so is the distance.
//    The K smallest (distance, B index) keys are the answer; no re-rank can
//    change them.
//  * FP8-E4M3: both patches centred and scaled per patch by 448 / max|x-128|,
//    dot products de-scaled in FP32 (times the two inverse scales), the two
//    norms those of the same quantized values (so the approximate distance is
//    |a' - b'|^2 of the quantized patches).  The S smallest approximate keys
//    form a shortlist that is re-ranked with exact integer distances into the
//    K list.
//
// Keys are (squared distance << 32) | B index, so ties go to the lower B
// index, the order of the exact field.  The first list entry is the match.
//
// Block: 256 threads, 64 queries of one leaf bucket (two 32-row MFMA blocks)
// against tiles of 128 leaf points (one 32-column block per wavefront).
// Patches are staged in LDS as bytes, rows padded to whole dwords.
#include "annf_internal.h"
#include "annf_patch.h"

#include <type_traits>

namespace annf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Where k_leaf_mm writes a single pass's result directly: knn (K per
// query, optional), nn and dist (optional); nn == nullptr: keys_out instead.
struct FieldOut {
    int *knn, *nn;
    float *dist;
};

constexpr int kMQ = 64;                // queries per block
constexpr int kNP = 128;               // points per tile
constexpr int kThreads = 256;
constexpr int kDS = kNP + 4;           // distance tile row stride (dwords)
constexpr uint64_t kNoKey = ~0ull;

// x -> round-to-nearest E4M3 code of (x - 128) * scale, byte by byte; bytes
// outside `mask` become code 0.
__device__ __forceinline__ uint32_t fp8_word(uint32_t w, float scale, uint32_t mask) {
    const float a0 = ((float)(w & 255u) - 128.0f) * scale;
    const float a1 = ((float)((w >> 8) & 255u) - 128.0f) * scale;
    const float a2 = ((float)((w >> 16) & 255u) - 128.0f) * scale;
    const float a3 = ((float)(w >> 24) - 128.0f) * scale;
    int packed = __builtin_amdgcn_cvt_pk_fp8_f32(a0, a1, 0, false);
    packed = __builtin_amdgcn_cvt_pk_fp8_f32(a2, a3, packed, true);
    return (uint32_t)packed & mask;
}

typedef float floatx8 __attribute__((ext_vector_type(8)));

// eight bytes (two dwords at p)
__device__ __forceinline__ uint2 raw8(const uint8_t *p) {
    return make_uint2(*reinterpret_cast<const uint32_t *>(p), *reinterpret_cast<const uint32_t *>(p + 4));
}

// eight bytes as BF16 (exact: bytes are small integers; the FP32 -> BF16
// step is one packed conversion per pair)
__device__ __forceinline__ bf16x8 bf16_of(uint2 w) {
    const floatx8 f = {(float)(w.x & 255u), (float)((w.x >> 8) & 255u), (float)((w.x >> 16) & 255u),
                       (float)(w.x >> 24), (float)(w.y & 255u), (float)((w.y >> 8) & 255u),
                       (float)((w.y >> 16) & 255u), (float)(w.y >> 24)};
    return __builtin_convertvector(f, bf16x8);
}

__device__ __forceinline__ long frag_fp8(const uint8_t *p) {
    const uint32_t w0 = *reinterpret_cast<const uint32_t *>(p);
    const uint32_t w1 = *reinterpret_cast<const uint32_t *>(p + 4);
    return (long)(((uint64_t)w1 << 32) | w0);
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const int lo = __shfl_xor((int)(uint32_t)v, m, 64);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Sorted key list of LM slots whose live part is the top L slots; the front
// slots hold 0, which never moves (an insert lands after every key <= it),
// so every index is static.
template <int LM>
__device__ __forceinline__ void klist_init(uint64_t (&l)[LM], int L) {
#pragma unroll
    for (int s = 0; s < LM; ++s) l[s] = (s < LM - L) ? 0ull : kNoKey;
}

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

// The same sorted list of 32-bit keys.  Short lists (LM <= 8) as a min/max
// network, applied to every key: slot s of the list with `key` added is the
// median of (old l[s-1], key, old l[s]), a key above every entry changes
// nothing.  Longer lists only when the key beats the last entry (most keys
// do not, and the network's cost grows with LM).
template <int LM>
__device__ __forceinline__ void klist_insert32(uint32_t (&l)[LM], uint32_t key) {
    if constexpr (LM <= 8) {
#pragma unroll
        for (int s = LM - 1; s >= 1; --s) l[s] = min(max(l[s - 1], key), l[s]);
        l[0] = min(l[0], key);
    } else if (key < l[LM - 1]) {
        bool le[LM];
#pragma unroll
        for (int s = 0; s < LM; ++s) le[s] = l[s] <= key;
#pragma unroll
        for (int s = LM - 1; s >= 1; --s) l[s] = le[s] ? l[s] : (le[s - 1] ? key : l[s - 1]);
        l[0] = le[0] ? l[0] : key;
    }
}

constexpr int kMaxRW = 10;   // dwords per padded patch row (p <= 12)

// Row r of a patch as rw (<= kMaxRW) dwords, tail masked, into w[]; every
// index is static (the loop runs kMaxRW times, guarded by rw).
__device__ __forceinline__ void row_words(const PatchGeom &g, long long off, int rw, uint32_t tmask,
                                          uint32_t (&w)[kMaxRW]) {
    const long long al = off & ~3LL;
    const int sh = (int)(off & 3);
    if (al + 4LL * (rw + 1) <= g.nbytes) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(g.base + al);
        uint32_t prev = src[0];
#pragma unroll
        for (int k = 0; k < kMaxRW; ++k) {
            if (k < rw) {
                const uint32_t next = src[k + 1];
                w[k] = __builtin_amdgcn_alignbyte(next, prev, sh);
                prev = next;
            } else {
                w[k] = 0u;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kMaxRW; ++k) {
            uint32_t v = 0;
            if (k < rw)
                for (int b = 0; b < 4; ++b) {
                    const long long o = off + 4 * k + b;
                    if (o < g.nbytes) v |= (uint32_t)g.base[o] << (8 * b);
                }
            w[k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < kMaxRW; ++k)
        if (k == rw - 1) w[k] &= tmask;
}

// Exact squared distances of A patch i to the B patches j[0..nc) (NC >= nc,
// static), row by row: each A row is read once for all candidates.
template <int NC>
__device__ __forceinline__ void exact_d2_rows(const PatchGeom &ga, long long i, const PatchGeom &gb,
                                              const uint32_t (&j)[NC], int nc, int rw, uint32_t tmask,
                                              uint32_t (&out)[NC]) {
    uint32_t saa = 0, sab[NC], sbb[NC];
    const long long oa = patch_origin(ga, i);
#pragma unroll
    for (int c = 0; c < NC; ++c) sab[c] = sbb[c] = 0u;
    for (int r = 0; r < ga.rows; ++r) {
        uint32_t wa[kMaxRW];
        row_words(ga, oa + r * ga.rstride, rw, tmask, wa);
#pragma unroll
        for (int k = 0; k < kMaxRW; ++k) saa = __builtin_amdgcn_udot4(wa[k], wa[k], saa, false);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (c >= nc) break;
            uint32_t wb[kMaxRW];
            row_words(gb, patch_origin(gb, j[c]) + r * gb.rstride, rw, tmask, wb);
#pragma unroll
            for (int k = 0; k < kMaxRW; ++k) {
                sbb[c] = __builtin_amdgcn_udot4(wb[k], wb[k], sbb[c], false);
                sab[c] = __builtin_amdgcn_udot4(wa[k], wb[k], sab[c], false);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = saa + sbb[c] - 2u * sab[c];
}

// Patch staging: fetch_rows / for_row_words, a quad of lanes per patch.
// Writes the fetched rows into the LDS patch row (dwords rows*rw..dkw zero)
// and returns this lane's share of sum(x^2) and sum(x).
template <int RL, int LW>
__device__ __forceinline__ void store_patch(const PatchGeom &g, long long o, const uint32_t (&w)[RL][LW],
                                            uint32_t *row, int rw, uint32_t tmask, int dkw, int q4, uint32_t &ss,
                                            uint32_t &sum) {
    ss = 0;
    sum = 0;
    for_row_words(g, o, w, q4, 4, rw, tmask, [&](int r, int k, uint32_t v) {
        row[r * rw + k] = v;
        ss = __builtin_amdgcn_udot4(v, v, ss, false);
        sum = __builtin_amdgcn_udot4(v, 0x01010101u, sum, false);
    });
    for (int k = g.rows * rw + q4; k < dkw; k += 4) row[k] = 0;
}

// In-place FP8 conversion of the rows this lane staged (after the quad knows
// the patch scale); returns this lane's share of the sum of squares of the
// converted values (in the scaled domain), so that the distance
// |a|^2 + |b|^2 - 2 a.b is formed from one consistent set of quantized values
// (a patch against itself then scores 0, whatever the rounding).
__device__ __forceinline__ float fp8_rows(uint32_t *row, int rows, int rw, uint32_t tmask, float scale, int q4) {
    float ss = 0.0f;
    for (int r = q4; r < rows; r += 4)
        for (int k = 0; k < rw; ++k) {
            const uint32_t w = fp8_word(row[r * rw + k], scale, k == rw - 1 ? tmask : ~0u);
            row[r * rw + k] = w;
            const float f0 = __builtin_amdgcn_cvt_f32_fp8((int)w, 0), f1 = __builtin_amdgcn_cvt_f32_fp8((int)w, 1);
            const float f2 = __builtin_amdgcn_cvt_f32_fp8((int)w, 2), f3 = __builtin_amdgcn_cvt_f32_fp8((int)w, 3);
            ss += (f0 * f0 + f1 * f1) + (f2 * f2 + f3 * f3);
        }
    return ss;
}

__device__ __forceinline__ float quad_sumf(float v) {
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    return v;
}

// max |x - 128| over the real bytes of the rows this lane staged
__device__ __forceinline__ uint32_t maxabs_rows(const uint32_t *row, int rows, int rw, uint32_t tmask, int q4) {
    int mx = 0, mn = 255;
    for (int r = q4; r < rows; r += 4)
        for (int k = 0; k < rw; ++k) {
            const uint32_t w = row[r * rw + k];
            const uint32_t m = k == rw - 1 ? tmask : ~0u;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if ((m >> (8 * b)) & 255u) {
                    const int x = (int)((w >> (8 * b)) & 255u);
                    mx = max(mx, x);
                    mn = min(mn, x);
                }
            }
        }
    if (mx < mn) return 0;   // this lane staged no bytes
    return (uint32_t)max(mx - 128, 128 - mn);
}

__device__ __forceinline__ uint32_t quad_sum(uint32_t v) {
    v += (uint32_t)__shfl_xor((int)v, 1, 64);
    v += (uint32_t)__shfl_xor((int)v, 2, 64);
    return v;
}

__device__ __forceinline__ uint32_t quad_max(uint32_t v) {
    v = max(v, (uint32_t)__shfl_xor((int)v, 1, 64));
    v = max(v, (uint32_t)__shfl_xor((int)v, 2, 64));
    return v;
}

__device__ __forceinline__ float sqrt_f(uint32_t d2) {
    // correctly rounded: the FP64 root rounded once more to FP32 (2*24+2 <= 53)
    return (float)__builtin_sqrt((double)d2);
}

template <int DK, int LM, bool FP8>
__global__ __launch_bounds__(kThreads, (FP8 && LM > 16) ? 1 : 2) void k_leaf_mm(TreeView t, PatchGeom ga, PatchGeom gb,
                                                         const int *__restrict__ pts_order,
                                                         const int *__restrict__ order,
                                                         const int *__restrict__ chunk_off,
                                                         const int4 *__restrict__ blk_info, int nprobe, int K,
                                                         int slot_div, uint64_t *__restrict__ keys_out,
                                                         FieldOut fo) {
    constexpr int RB = DK + 4;                              // LDS patch stride (bytes, odd dword count)
    constexpr int kBD = (kNP * RB > kMQ * kDS * 4) ? kNP * RB : kMQ * kDS * 4;
    extern __shared__ __align__(16) uint8_t smem[];
    uint8_t *Aq = smem;                                      // kMQ x RB
    uint8_t *BD = Aq + kMQ * RB;                             // points, then distances, then candidates
    uint32_t *Dm = reinterpret_cast<uint32_t *>(BD);         // kMQ x kDS
    uint64_t *cand = reinterpret_cast<uint64_t *>(BD);       // kMQ x LM
    uint32_t *na = reinterpret_cast<uint32_t *>(BD + kBD);   // kMQ
    uint32_t *nb = na + kMQ;                                 // kNP
    int *nbsum = reinterpret_cast<int *>(nb + kNP);          // kNP
    int *qid = nbsum + kNP;                                  // kMQ
    int *borig = qid + kMQ;                                  // kNP
    float *sa = reinterpret_cast<float *>(borig + kNP);      // kMQ
    float *sb = sa + kMQ;                                    // kNP

    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int q = tid >> 2, q4 = tid & 3;                    // query row / quad lane
    const int nleaves = 1 << t.depth;
    const int rw = (ga.rowlen + 3) >> 2;
    const int tail = ga.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : ~0u;
    const int D = ga.rows * ga.rowlen;
    constexpr int dkw = DK / 4;
    const int b = blockIdx.x;
    // this block's chunk (blk_info holds gridDim.x pairs, the used ones set)
    const int4 bi0 = blk_info[2 * b], bi1 = blk_info[2 * b + 1];
    if (b >= chunk_off[nleaves]) return;
    const int leaf = bi0.x, qb = bi0.y, nqb = bi0.z;
    if (nqb <= 0) return;

    using SS = StageShape<DK>;
    constexpr int RL = SS::RL, LW = SS::LW;
    constexpr int kBP = kNP * 4 / kThreads;   // B patches per quad lane and tile

    // Tiles of up to kNP points of the probed leaves, in order.  The first
    // tile's indices go out before the query rows and its points' rows right
    // after them, so a block waits on global memory about three times.
    int pr = 0, t0 = 0;
    int from = bi0.w, cnt = bi1.x;
    auto seek = [&]() -> bool {   // the next non-empty tile at or after (pr, t0)
        while (t0 >= cnt) {
            if (++pr >= nprobe) return false;
            const int pl = leaf ^ pr;
            from = t.leaves[2 * pl];
            cnt = t.leaves[2 * pl + 1] - from;
            t0 = 0;
        }
        return true;
    };
    int np = 0, bj[kBP];
    long long ob[kBP];
    uint32_t rows_b[kBP][RL][LW];
    auto tile_index = [&]() {
        np = min(kNP, cnt - t0);
#pragma unroll
        for (int u = 0; u < kBP; ++u) {
            const int i = (tid >> 2) + u * (kThreads / 4);
            bj[u] = i < np ? pts_order[from + t0 + i] : 0;
        }
    };
    auto tile_fetch = [&]() {
#pragma unroll
        for (int u = 0; u < kBP; ++u) {
            ob[u] = patch_origin(gb, bj[u]);
            fetch_rows(gb, ob[u], (tid >> 2) + u * (kThreads / 4) < np, q4, 4, rows_b[u]);
        }
    };

    // queries: one quad per query; stored with the first tile's points
    const bool qv = q < nqb;
    const int slot = qv ? order[qb + q] : 0;   // output slot; its A patch is slot / slot_div
    bool more = seek();
    if (more) tile_index();
    const long long oa = patch_origin(ga, slot / slot_div);
    uint32_t rows_a[RL][LW];
    fetch_rows(ga, oa, qv, q4, 4, rows_a);
    if (q4 == 0) qid[q] = slot;
    if (more) tile_fetch();
    bool a_staged = false;

    // this lane's LM smallest keys (every slot live)
    uint64_t lst[LM];
    klist_init(lst, LM);
    bool first_tile = true;

    while (more) {
        __syncthreads();   // the previous tile's distances are consumed
        if (!a_staged) {
            a_staged = true;
            uint32_t *row = reinterpret_cast<uint32_t *>(Aq + q * RB);
            uint32_t ss, sm;
            store_patch(ga, oa, rows_a, row, rw, tmask, dkw, q4, ss, sm);
            ss = quad_sum(ss);
            sm = quad_sum(sm);
            if constexpr (FP8) {
                const uint32_t mx = quad_max(maxabs_rows(row, ga.rows, rw, tmask, q4));
                const float sc = mx ? 448.0f / (float)mx : 1.0f;
                const float inv = mx ? (float)mx / 448.0f : 1.0f;
                const float qq = quad_sumf(fp8_rows(row, ga.rows, rw, tmask, sc, q4));
                if (q4 == 0) {
                    sa[q] = inv;   // the inverse scale
                    // |a - 128|^2 of the quantized patch, de-scaled
                    na[q] = __float_as_uint(qq * (inv * inv));
                }
            } else {
                if (q4 == 0) na[q] = ss;
            }
        }
        const int tnp = np;   // this tile's point count
#pragma unroll
        for (int u = 0; u < kBP; ++u) {
            const int i = (tid >> 2) + u * (kThreads / 4);
            uint32_t *row = reinterpret_cast<uint32_t *>(BD + i * RB);
            uint32_t ss, sm;
            store_patch(gb, ob[u], rows_b[u], row, rw, tmask, dkw, q4, ss, sm);
            ss = quad_sum(ss);
            sm = quad_sum(sm);
            if constexpr (FP8) {
                const uint32_t mx = quad_max(maxabs_rows(row, gb.rows, rw, tmask, q4));
                const float sc = mx ? 448.0f / (float)mx : 1.0f;
                const float inv = mx ? (float)mx / 448.0f : 1.0f;
                const float qq = quad_sumf(fp8_rows(row, gb.rows, rw, tmask, sc, q4));
                if (q4 == 0) {
                    sb[i] = inv;
                    nb[i] = __float_as_uint(qq * (inv * inv));
                }
            } else {
                if (q4 == 0) nb[i] = ss;
            }
            if (q4 == 0) borig[i] = bj[u];
        }
        __syncthreads();
        {
            // 64 x 32 block of dot products per wavefront
            // acc0 / acc1: query rows 0-31 / 32-63; BF16 splits each over
            // even (x0) and odd (x1) k-steps
            floatx16 acc0 = {}, acc1 = {}, acc0x = {}, acc1x = {};
            {
                const uint8_t *ar0 = Aq + (lane & 31) * RB + 8 * (lane >> 5);
                const uint8_t *ar1 = ar0 + 32 * RB;
                const uint8_t *br = BD + (32 * wid + (lane & 31)) * RB + 8 * (lane >> 5);
                if constexpr (FP8) {
#pragma unroll 5
                    for (int k0 = 0; k0 < DK; k0 += 16) {
                        const long fb = frag_fp8(br + k0);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(frag_fp8(ar0 + k0), fb, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(frag_fp8(ar1 + k0), fb, acc1, 0, 0, 0);
                    }
                } else {
                    // the next 16-deep step's bytes are read before this step's
                    // are converted, so each LDS read has a step to land
                    uint2 wa0 = raw8(ar0), wa1 = raw8(ar1), wb = raw8(br);
#pragma unroll
                    for (int k0 = 0; k0 < DK; k0 += 16) {
                        uint2 na0 = wa0, na1 = wa1, nbw = wb;
                        if (k0 + 16 < DK) {
                            na0 = raw8(ar0 + k0 + 16);
                            na1 = raw8(ar1 + k0 + 16);
                            nbw = raw8(br + k0 + 16);
                        }
                        const bf16x8 fb = bf16_of(wb);
                        if ((k0 & 16) == 0) {
                            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf16_of(wa0), fb, acc0, 0, 0, 0);
                            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf16_of(wa1), fb, acc1, 0, 0, 0);
                        } else {
                            acc0x = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf16_of(wa0), fb, acc0x, 0, 0, 0);
                            acc1x = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf16_of(wa1), fb, acc1x, 0, 0, 0);
                        }
                        wa0 = na0;
                        wa1 = na1;
                        wb = nbw;
                    }
                }
            }
            const int pcol = 32 * wid + (lane & 31);
            const uint32_t nbp = nb[pcol];
            const float sbp = sb[pcol];
            __syncthreads();   // every wavefront is done with the staged points
            {
                const bool pv = pcol < tnp;
                // a column past the tile's points makes every key 0xFFFFFFFF
                const uint32_t colbits = pv ? (uint32_t)pcol : 0xFFFFFFFFu;
#pragma unroll
                for (int r4 = 0; r4 < 4; ++r4) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        // the accumulator rows r4 * 4 + 0..3 are four consecutive queries
                        const int q0 = r4 * 8 + (lane >> 5) * 4 + 32 * h;
                        const uint4 na4 = *reinterpret_cast<const uint4 *>(na + q0);
                        float4 sa4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if constexpr (FP8) sa4 = *reinterpret_cast<const float4 *>(sa + q0);
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const int r = 4 * r4 + i, qq = q0 + i;
                            const uint32_t naq = i == 0 ? na4.x : (i == 1 ? na4.y : (i == 2 ? na4.z : na4.w));
                            const float acc = h ? acc1[r] : acc0[r];
                            uint32_t d;
                            if constexpr (FP8) {
                                const float saq = i == 0 ? sa4.x : (i == 1 ? sa4.y : (i == 2 ? sa4.z : sa4.w));
                                const float dd = (__uint_as_float(naq) + __uint_as_float(nbp)) -
                                                 2.0f * (acc * (saq * sbp));
                                d = dd <= 0.0f ? 0u : (dd >= 4.0e9f ? 0xFFFFFFFEu : (uint32_t)dd);
                                d = pv ? d : 0xFFFFFFFFu;
                            } else {
                                // a.b as the sum of two exact integer halves; the key
                                // (d2 << 7) | tile position orders like (d2, B index)
                                // (positions follow ascending B index, d2 < 2^25)
                                const float accx = h ? acc1x[r] : acc0x[r];
                                d = ((naq + nbp - 2u * ((uint32_t)(int)acc + (uint32_t)(int)accx)) << 7) | colbits;
                            }
                            Dm[qq * kDS + pcol] = d;
                        }
                    }
                }
            }
            __syncthreads();
            if (q < nqb) {
                const uint32_t *dr = Dm + q * kDS;
                if constexpr (FP8) {
                    for (int p = q4; p < tnp; p += 4) {
                        const uint64_t key = ((uint64_t)dr[p] << 32) | (uint32_t)borig[p];
                        if (key < lst[LM - 1]) klist_insert(lst, key);
                    }
                } else {
                    // this tile's best with 32-bit keys, then as (d2, B index) keys
                    uint32_t l32[LM];
#pragma unroll
                    for (int s2 = 0; s2 < LM; ++s2) l32[s2] = 0xFFFFFFFFu;
                    // all kNP columns of the row are written (0xFFFFFFFF past
                    // the tile's points): eight reads in flight, then their inserts
#pragma unroll
                    for (int p0 = q4; p0 < kNP; p0 += 32) {
                        uint32_t kk[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) kk[j] = dr[p0 + 4 * j];
#pragma unroll
                        for (int j = 0; j < 8; ++j) klist_insert32(l32, kk[j]);
                    }
                    // the B indices of the tile's best, every read in flight first
                    // (k & 127 is a tile position, also for the empty key)
                    int bo[LM];
#pragma unroll
                    for (int s2 = 0; s2 < LM; ++s2) bo[s2] = borig[l32[s2] & 127u];
#pragma unroll
                    for (int s2 = 0; s2 < LM; ++s2) {
                        const uint32_t k = l32[s2];
                        const uint64_t key =
                            k == 0xFFFFFFFFu ? kNoKey : (((uint64_t)(k >> 7) << 32) | (uint32_t)bo[s2]);
                        if (first_tile) {
                            lst[s2] = key;   // ascending already
                        } else if (key < lst[LM - 1]) {
                            klist_insert(lst, key);
                        }
                    }
                }
            }
            first_tile = false;
        }
        // the next tile's loads, in flight across the barrier
        t0 += kNP;
        more = seek();
        if (more) {
            tile_index();
            tile_fetch();
        }
    }
    __syncthreads();   // the distance tile is free for the candidate lists

    // merge the quad's four lists: the LM smallest keys, in order, into cand[q]
    for (int s = 0; s < LM; ++s) {
        const uint64_t head = lst[0];
        uint64_t m = head;
        const uint64_t m1 = shfl_xor64(m, 1);
        m = m1 < m ? m1 : m;
        const uint64_t m2 = shfl_xor64(m, 2);
        m = m2 < m ? m2 : m;
        if (head == m && m != kNoKey) {
#pragma unroll
            for (int k = 0; k < LM - 1; ++k) lst[k] = lst[k + 1];
            lst[LM - 1] = kNoKey;
        }
        if (q4 == 0) cand[q * LM + s] = m;
    }
    if constexpr (FP8) {
        __syncthreads();
        // exact re-rank of the shortlist
        if (q < nqb) {
            // this lane's entries q4, q4 + 4, ... of the shortlist
            constexpr int NC = (LM + 3) / 4;
            uint32_t jj[NC], d2[NC];
            int nc = 0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const int s = q4 + 4 * c;
                const uint64_t key = s < LM ? cand[q * LM + s] : kNoKey;
                jj[c] = (uint32_t)key;
                if (key != kNoKey) nc = c + 1;   // live entries come first
            }
            exact_d2_rows(ga, qid[q] / slot_div, gb, jj, nc, rw, tmask, d2);
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c < nc) cand[q * LM + q4 + 4 * c] = ((uint64_t)d2[c] << 32) | jj[c];
        }
        __syncthreads();
        if (q < nqb && q4 == 0) {
            // K smallest exact keys, in order, into the front of cand[q]
            uint64_t best[LM];
            klist_init(best, K);
            for (int s = 0; s < LM; ++s) {
                const uint64_t key = cand[q * LM + s];
                if (key < best[LM - 1]) klist_insert(best, key);
            }
#pragma unroll
            for (int s = 0; s < LM; ++s)
                if (s >= LM - K) cand[q * LM + s - (LM - K)] = best[s];
        }
    }
    __syncthreads();
    if (q < nqb) {
        const size_t o = (size_t)qid[q] * K;
        if (fo.nn) {
            // the final field (what k_finish writes from a single pass)
            for (int s = q4; s < K; s += 4) {
                const uint64_t key = cand[q * LM + s];
                const int j = key == kNoKey ? -1 : (int)(uint32_t)key;
                if (fo.knn) fo.knn[o + s] = j;
                if (s == 0) {
                    fo.nn[qid[q]] = j;
                    if (fo.dist) fo.dist[qid[q]] = key == kNoKey ? kFltMax : sqrt_f((uint32_t)(key >> 32));
                }
            }
        } else {
            for (int s = q4; s < K; s += 4) keys_out[o + s] = cand[q * LM + s];
        }
    }
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

// Block c's work record: the leaf l with chunk_off[l] <= c < chunk_off[l + 1]
// (binary search, one thread per block, so a leaf with many chunks is no
// serial tail); blk_info[2c] = (leaf, first query slot in `order`, queries,
// first point of the leaf), blk_info[2c + 1].x = points of the leaf.
__global__ void k_block_info(const int *__restrict__ chunk_off, const int *__restrict__ offset,
                             const int *__restrict__ leaves, int nleaves, int nblk, int4 *__restrict__ blk_info) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nblk || c >= chunk_off[nleaves]) return;
    int lo = 0, hi = nleaves - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (chunk_off[m] <= c) lo = m; else hi = m - 1;
    }
    const int l = lo;
    const int qb = offset[l] + (c - chunk_off[l]) * kMQ, from = leaves[2 * l];
    blk_info[2 * c] = make_int4(l, qb, min(kMQ, offset[l + 1] - qb), from);
    blk_info[2 * c + 1] = make_int4(leaves[2 * l + 1] - from, 0, 0, 0);
}

__global__ void k_fill_keys(uint64_t *__restrict__ keys, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = kNoKey;
}

// The K smallest of a query's first-pass keys and its propagated pairs'
// keys (k_prop_pairs leaves out every leaf the first pass scanned and
// repeats, so the leaves are disjoint and no B index comes twice) -> knn,
// nn, dist.
template <int KM>
__global__ void k_finish(const uint64_t *__restrict__ keys1, const uint64_t *__restrict__ keys2, int q0, int nq,
                         int K, int prop_k, int *__restrict__ knn, int *__restrict__ nn, float *__restrict__ dist,
                         uint64_t *__restrict__ keys_out) {
    const int ql = blockIdx.x * blockDim.x + threadIdx.x;
    if (ql >= nq) return;
    const int q = q0 + ql;
    uint64_t l[KM];
    klist_init(l, K);
    for (int s = 0; s < K; ++s) {
        const uint64_t key = keys1[(size_t)q * K + s];
        if (key < l[KM - 1]) klist_insert(l, key);
    }
    if (keys2) {
        const uint64_t *p2 = keys2 + (size_t)ql * prop_k * K;
        for (int s = 0; s < prop_k * K; ++s) {
            const uint64_t key = p2[s];
            if (key < l[KM - 1]) klist_insert(l, key);
        }
    }
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

template <typename Fn>
hipError_t with_dk_lm(int dk_need, int lm_need, Fn &&fn) {
    auto pick_lm = [&](auto dk) -> hipError_t {
        if (lm_need <= 5) return fn(dk, std::integral_constant<int, 5>{});
        if (lm_need <= 8) return fn(dk, std::integral_constant<int, 8>{});
        if (lm_need <= 16) return fn(dk, std::integral_constant<int, 16>{});
        if (lm_need <= 20) return fn(dk, std::integral_constant<int, 20>{});
        return fn(dk, std::integral_constant<int, 32>{});
    };
    if (dk_need <= 176) return pick_lm(std::integral_constant<int, 176>{});
    if (dk_need <= 256) return pick_lm(std::integral_constant<int, 256>{});
    if (dk_need <= 400) return pick_lm(std::integral_constant<int, 400>{});
    return pick_lm(std::integral_constant<int, 512>{});
}

}  // namespace

int leafmm_chunk() { return kMQ; }

// padded patch length (bytes) the kernel stages: rows of whole dwords,
// rounded up to the MFMA depth of 16
int leafmm_dk(int p) {
    const int rw = (3 * p + 3) / 4;
    return (p * rw * 4 + 15) / 16 * 16;
}

bool leafmm_supported(int p) { return p >= 1 && leafmm_dk(p) <= 512; }

hipError_t prop_pairs(const uint64_t *keys1, int q0, int nq, int K, int prop_k, int nprobe, const int *natleaf,
                      const int *pli, int n_cols, int n_b, int *pairleaf, int *count, hipStream_t st) {
    const long long n = (long long)nq * prop_k;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prop_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys1, q0, nq, K, prop_k,
                       nprobe, natleaf, pli, n_cols, n_b, pairleaf, count);
    return hipGetLastError();
}

hipError_t fill_keys(uint64_t *keys, long long n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, keys, n);
    return hipGetLastError();
}

hipError_t finish_field(const uint64_t *keys1, const uint64_t *keys2, int q0, int nq, int K, int prop_k, int *knn,
                        int *nn, float *dist, uint64_t *keys_out, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    const dim3 g((unsigned)((nq + 255) / 256)), b(256);
    if (K <= 8)
        hipLaunchKernelGGL(k_finish<8>, g, b, 0, st, keys1, keys2, q0, nq, K, prop_k, knn, nn, dist, keys_out);
    else
        hipLaunchKernelGGL(k_finish<32>, g, b, 0, st, keys1, keys2, q0, nq, K, prop_k, knn, nn, dist, keys_out);
    return hipGetLastError();
}

static size_t leafmm_blocks(int nq, int nleaves) { return (size_t)nleaves + ((size_t)nq + kMQ - 1) / kMQ; }

size_t leafmm_scratch_bytes(int nq, int nleaves) { return leafmm_blocks(nq, nleaves) * 2 * sizeof(int4); }

hipError_t leaf_point_order(const TreeView &t, int *out, hipStream_t st) {
    hipLaunchKernelGGL(k_leaf_order, dim3(1u << t.depth), dim3(kLeafSortTPB), 0, st, t.leaves, t.orig, out);
    return hipGetLastError();
}

hipError_t leaf_mm(const TreeView &t, const PatchGeom &ga, const PatchGeom &gb, const int *pts_order, const int *order,
                   const int *offset, const int *chunk_off, void *scratch, int nq, int nprobe, int shortlist, int K,
                   int fp8, int slot_div, uint64_t *keys_out, int *knn, int *nn, float *dist, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    const int p = ga.rows;
    if (!leafmm_supported(p) || gb.rows != p || ga.rowlen != 3 * p || gb.rowlen != 3 * p) return hipErrorInvalidValue;
    const int nleaves = 1 << t.depth;
    if (nprobe < 1 || nprobe > nleaves || K < 1 || K > 32 || slot_div < 1) return hipErrorInvalidValue;
    // list length: K for BF16 (exact keys), the shortlist for FP8; rounded
    // up to an instantiated length (a longer list only adds candidates)
    const int L = fp8 ? shortlist : K;
    if (L < K || L > 32) return hipErrorInvalidValue;
    const unsigned nblk = (unsigned)leafmm_blocks(nq, nleaves);
    int4 *blk_info = static_cast<int4 *>(scratch);
    hipLaunchKernelGGL(k_block_info, dim3((nblk + 255) / 256), dim3(256), 0, st, chunk_off, offset, t.leaves,
                       nleaves, (int)nblk, blk_info);
    return with_dk_lm(leafmm_dk(p), L, [&](auto dk, auto lm) -> hipError_t {
        constexpr int DK = decltype(dk)::value;
        constexpr int LM = decltype(lm)::value;
        constexpr int RB = DK + 4;
        constexpr int kBD = (kNP * RB > kMQ * kDS * 4) ? kNP * RB : kMQ * kDS * 4;
        static_assert(kBD >= kMQ * LM * 8, "candidate lists must fit the tile region");
        const size_t lds = (size_t)kMQ * RB + kBD + (size_t)(kMQ + kNP + kNP + kMQ + kNP + kMQ + kNP) * 4;
        auto launch = [&](auto f8) -> hipError_t {
            constexpr bool F8 = decltype(f8)::value;
            hipError_t e = hipFuncSetAttribute((const void *)k_leaf_mm<DK, LM, F8>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_leaf_mm<DK, LM, F8>), dim3(nblk), dim3(kThreads), lds, st, t, ga, gb,
                               F8 ? t.orig : pts_order, order,
                               chunk_off, blk_info, nprobe, K, slot_div, keys_out, FieldOut{knn, nn, dist});
            return hipGetLastError();
        };
        return fp8 ? launch(std::true_type{}) : launch(std::false_type{});
    });
}

}  // namespace annf
