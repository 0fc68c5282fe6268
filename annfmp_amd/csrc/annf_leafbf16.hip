// Exact full-patch leaf k-NN with BF16 operands on MFMA (the BF16 precision
// variant and the learned-encoder path).
//
// Arithmetic.  Both images are first rewritten once per call as BF16 values
// of x - 128 (k_center_bf16), which BF16 holds exactly, so staging a patch
// is a copy.  For a query patch a' and a leaf point b' (centred), every
// product is an integer of magnitude <= 16384 and every partial sum of a
// dot product over D <= 432 elements stays below 2^23, so an FP32
// accumulator holds a'.b' exactly whatever the MFMA's summation order (two
// accumulators, even and odd k-steps, sum below 2^24: also exact), and
//     d2 = |a'|^2 + |b'|^2 - 2 a'.b' = |a - b|^2
// is the exact integer squared distance.  The norms |x'|^2 are exact
// integer sums per patch (k_patch_norms, a p x p box sum of per-pixel
// squares).
//
// Selection keys.  For one query, |a'|^2 is a constant, so within the scan
// a candidate is ranked by v = |b'|^2 + OFF - 2 a'.b' (OFF = 16384 D makes
// v >= 0 and v <= 65536 D); the 32-bit key is (v << PB) | position, with PB
// the position bits that fit (7 for D <= 512 / 8 for D <= 255), so keys
// order like (d2, position).  A staged point carries base = ((|b'|^2 + OFF)
// << PB) | position, and key = base - (2 << PB) * a'.b' is one
// v_mad_i32_i24 after the conversion of the accumulator; a point past the
// tile's end has base = 0xFFFFFFFF and zero operands, so its key is the
// empty key.  Positions follow ascending B index within a leaf (the points
// are staged in that order), so ties go to the lower B index.  Every
// 2^PB-point segment of a leaf, each query's K smallest 32-bit keys become
// 64-bit keys (d2 << 32) | B index, merged across segments and probed
// leaves: the K smallest (distance, B index), the exact field's order.
//
// Layout.  A patch row of 3p elements is read as G = ceil(3p / 8) groups of
// 8 consecutive elements (16 bytes); elements of the last group past the
// row belong to the next pixels and are zeroed in the query operand, so they
// add nothing (norms are exact separately).  Group g = r G + j is row r,
// elements 8 j .. 8 j + 7; zeros pad the group count to an even number
// 2 ks.  k-step s (16 deep) takes groups 2s and 2s + 1, so the two lanes
// of a query read neighbouring pieces of a row.  Each image is kept twice as centred
// BF16, the second copy shifted by one element, so every group is one
// 4-byte-aligned 16-byte load from one of the copies.  p = 11: 55 groups,
// 28 k-steps.
//
// Block: 256 threads; 128 queries of one leaf bucket, 32 per wavefront,
// held in registers as MFMA B operands (lane l: query l & 31, groups
// 2s + (l >> 5)); the leaf's points are staged 32 at a time in LDS in
// fragment order ([group][point][8 elements], groups 528 bytes apart), read
// as MFMA A operands; eight consecutive threads stage eight consecutive
// groups of one point, so a load instruction touches few cache lines.  A segment's point indices are read into LDS once; the next
// tile's global loads are issued before the current tile's MFMAs, into
// registers, and written to LDS after them.  A
// 32x32 output block leaves each lane 16 distances of its own query, so
// selection runs in registers (a v_med3 network per key) with no LDS round
// trip; the two lanes of a query merge their lists at the end of a segment.
#include "annf_internal.h"
#include "annf_lists.h"

#include <type_traits>

namespace annf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4a __attribute__((ext_vector_type(4), aligned(4)));

constexpr int kQ = 128;          // queries per block (4 wavefronts x 32)
constexpr int kTP = 32;          // points per staged tile (one 32x32 MFMA block)
constexpr int kThreads = 256;
constexpr int kGroupSets = kThreads / kTP;   // staging threads per point
constexpr int kSlot = (kTP + 1) * 16;        // bytes between groups in LDS (a bank offset per group)
constexpr int kMaxGroups = 60;   // groups of a patch (p <= 12: 12 rows x 5)
constexpr int kMaxSeg = 256;     // positions per segment (2^PB)
constexpr uint64_t kNoKey = ~0ull;

// A centred BF16 image twice: px0[e] is element e = (y * width + x) * 3 +
// channel (BF16 bits of byte - 128), px1[e] is element e + 1.  Both hold at
// least 16 zero elements past the image, so a group read past a row's end
// stays inside the buffer.
struct CImg {
    const uint16_t *px0;   // the first copy; the second starts npad elements later
    uint32_t npad;         // elements per copy
    int width;             // pixels per image row
    int ncols;             // patches per image row
};

// The group layout of a patch (see the file comment) for patch size p.
struct PatchPack {
    int p, G;       // rows, groups per row
    int ngroups;    // p * G
    int ks;         // k-steps: ceil(ngroups / 2)
    int pb;         // position bits of the 32-bit keys
    uint32_t off;   // OFF = 16384 D
};

// element offset of patch i (images of fewer than 2^31 elements)
__device__ __forceinline__ uint32_t patch_base(const CImg &im, int i) {
    const unsigned ui = (unsigned)i, nc = (unsigned)im.ncols;
    const unsigned y = ui / nc, x = ui - y * nc;
    return (y * (unsigned)im.width + x) * 3u;
}

// elements e .. e + 7: one 16-byte load from the copy in which e is even
// (element e - 1 of the second copy holds element e)
__device__ __forceinline__ u32x4 load_group(const CImg &im, uint32_t e) {
    const uint32_t odd = e & 1u;
    const uint32_t idx = (e - odd) + odd * im.npad;
    return *reinterpret_cast<const u32x4a *>(im.px0 + idx);
}

struct Out {
    int *knn, *nn;
    float *dist;
    uint64_t *keys;   // used when nn == nullptr
};

template <int KSM, int LM, bool EXACT>
__global__ __launch_bounds__(kThreads, 2) void k_leaf_bf16(TreeView t, CImg ia, CImg ib, PatchPack pk,
                                                         const uint32_t *__restrict__ norm_a,
                                                         const uint32_t *__restrict__ norm_b,
                                                         const int *__restrict__ pts_order,
                                                         const int *__restrict__ order,
                                                         const int *__restrict__ chunk_off,
                                                         const int4 *__restrict__ blk_info, int nprobe, int K,
                                                         int slot_div, int a0, Out out) {
    constexpr int NGM = 2 * KSM;
    constexpr int GPT = (NGM + kGroupSets - 1) / kGroupSets;   // groups per staging thread and tile
    __shared__ __align__(16) uint8_t bt[NGM * kSlot];        // [group][point][8 BF16]
    __shared__ __align__(16) uint32_t base[kTP];
    __shared__ __align__(16) u32x4 gmask[NGM];               // element masks of the groups
    __shared__ int borig[kMaxSeg];
    __shared__ int goff_a[NGM], goff_b[NGM];                 // group element offsets in A's / B's image

    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int h = lane >> 5, qc = lane & 31;
    const int nleaves = 1 << t.depth;
    const int b = blockIdx.x;
    if (b >= chunk_off[nleaves]) return;
    const int4 bi0 = blk_info[2 * b], bi1 = blk_info[2 * b + 1];
    const int leaf = bi0.x, qb = bi0.y, nqb = bi0.z;
    if (nqb <= 0) return;

    // group tables: offsets (clamped into the patch for the padding groups)
    // and the element masks of the staged points (0 for padding groups and
    // for elements past the row)
    for (int g = tid; g < NGM; g += kThreads) {
        const int gg = min(g, pk.ngroups - 1);
        const int r = gg / pk.G, j = gg - r * pk.G;
        goff_a[g] = r * 3 * ia.width + 8 * j;
        goff_b[g] = r * 3 * ib.width + 8 * j;
        const int valid = g < pk.ngroups ? min(8, 3 * pk.p - 8 * j) : 0;
        u32x4 m;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            m[k] = 2 * k + 1 < valid ? 0xFFFFFFFFu : (2 * k < valid ? 0x0000FFFFu : 0u);
        gmask[g] = m;
    }

    const int qi = 32 * wid + qc;
    const bool qv = qi < nqb;
    const bool wave_live = 32 * wid < nqb;
    int slot = 0;
    if (qv) slot = order[qb + qi];

    const int pb = pk.pb;
    const int seg = 1 << pb;
    const int mul = -(2 << pb);

    // tile walk: tiles of kTP points over the probed leaves; a segment is
    // up to 2^pb consecutive points of one leaf, its point indices in borig
    int pr = 0, from = bi0.w, cnt = bi1.x, t0 = 0;
    // this thread's staging slot: point sp, groups sg, sg + kGroupSets, ...
    const int sp = tid / kGroupSets, sg = tid % kGroupSets;
    u32x4 raw[GPT];
    bool rvalid = false;
    uint32_t rbase = 0xFFFFFFFFu;

    auto load_segment = [&]() {
        const int n = min(seg, cnt - t0);
        for (int i = tid; i < n; i += kThreads) borig[i] = pts_order[from + t0 + i];
    };
    // global loads of the tile at t0 (np points) into raw / rbase; every
    // address is inside the image (absent points read patch 0)
    auto fetch_tile = [&](int np) {
        const int segbase = t0 & ~(seg - 1);
        rvalid = sp < np;
        const int j = rvalid ? borig[t0 - segbase + sp] : 0;
        if (sg == 0)
            rbase = rvalid ? (((norm_b[j] + pk.off) << pb) | (uint32_t)(t0 - segbase + sp)) : 0xFFFFFFFFu;
        const uint32_t e0 = patch_base(ib, j);
#pragma unroll
        for (int i = 0; i < GPT; ++i) {
            const int g = min(sg + kGroupSets * i, NGM - 1);
            raw[i] = load_group(ib, e0 + (uint32_t)goff_b[g]);
        }
    };
    // raw -> LDS (after the previous tile is consumed), elements past a row
    // zeroed (gmask); absent points and padding groups as zeros
    auto store_tile = [&]() {
        if (sg == 0) base[sp] = rbase;
#pragma unroll
        for (int i = 0; i < GPT; ++i) {
            const int g = sg + kGroupSets * i;
            if (g < 2 * pk.ks) {
                const u32x4 v = rvalid ? (raw[i] & gmask[g]) : u32x4{0u, 0u, 0u, 0u};
                *reinterpret_cast<u32x4 *>(bt + g * kSlot + sp * 16) = v;
            }
        }
    };

    // the first segment's indices, then this lane's A fragments (groups
    // h ks + s) and the first tile, all loads in flight together
    load_segment();
    __syncthreads();
    // the query operand (groups 2s + h) as loaded: its elements past a row
    // (and its padding group) meet zeros in the staged points
    uint32_t na = 0;
    bf16x8 af[KSM];
    {
        const int apatch = a0 + slot / slot_div;
        if (qv) na = norm_a[apatch];
        const uint32_t e0 = patch_base(ia, apatch);
#pragma unroll
        for (int s = 0; s < KSM; ++s) {
            const int g = min(2 * s + h, NGM - 1);
            af[s] = __builtin_bit_cast(bf16x8, load_group(ia, e0 + (uint32_t)goff_a[g]));
        }
    }
    int np = min(kTP, cnt - t0);
    fetch_tile(np);

    uint32_t l32[LM];
    uint64_t l64[LM];
#pragma unroll
    for (int s = 0; s < LM; ++s) {
        l32[s] = s < LM - K ? 0u : 0xFFFFFFFFu;
        l64[s] = s < LM - K ? 0ull : kNoKey;
    }
    bool first_seg = true;

    for (;;) {
        // tile (t0, np) is in raw: into LDS
        store_tile();
        __syncthreads();
        const int segbase = t0 & ~(seg - 1);
        const int seg_end = min(segbase + seg, cnt);
        const int cur_np = np;
        // the next tile of this segment goes out now, in flight during the
        // compute (a new segment or leaf needs its indices first)
        const int nt0 = t0 + cur_np;
        const bool next_same_seg = nt0 < seg_end;
        if (next_same_seg) {
            t0 = nt0;
            np = min(kTP, cnt - t0);
            fetch_tile(np);
        }
        if (wave_live) {
            // even and odd k-steps accumulate separately (two independent
            // MFMA chains); each partial sum stays below 2^23 and their sum
            // below 2^24, so both and the total are exact
            floatx16 acc = {}, acc1 = {};
            const uint8_t *br = bt + h * kSlot + qc * 16;
            if constexpr (EXACT) {
                // fragment reads run kAhead k-steps in front of the MFMAs
                constexpr int kAhead = 4;
#pragma unroll
                for (int s = 0; s < KSM; ++s) {
                    const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(br + s * (2 * kSlot));
                    if (s & 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bv, af[s], acc1, 0, 0, 0);
                    else acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bv, af[s], acc, 0, 0, 0);
                }
                __builtin_amdgcn_sched_group_barrier(0x100, kAhead, 0);   // LDS reads
#pragma unroll
                for (int s = 0; s < KSM; ++s) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);     // MFMA
                    if (s + kAhead < KSM) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                }
            } else {
#pragma unroll
                for (int s = 0; s < KSM; ++s) {
                    if (s < pk.ks) {
                        const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(br + s * (2 * kSlot));
                        if (s & 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bv, af[s], acc1, 0, 0, 0);
                        else acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bv, af[s], acc, 0, 0, 0);
                    }
                }
            }
            acc += acc1;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                // rows 8 j + 4 h + 0..3 of the tile
                const u32x4 bq = *reinterpret_cast<const u32x4 *>(base + 8 * j + 4 * h);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int dot = (int)acc[4 * j + i];
                    const uint32_t key = bq[i] + (uint32_t)__mul24(dot, mul);
                    list_insert<LM>(l32, key);
                }
            }
        }
        if (!next_same_seg && wave_live) {
            // the segment's K best of this query: merge the two lanes' lists,
            // then as (d2, B index) keys into the running list
            uint32_t other[LM];
#pragma unroll
            for (int s = 0; s < LM; ++s) other[s] = (uint32_t)__shfl_xor((int)l32[s], 32, 64);
#pragma unroll
            for (int s = 0; s < LM; ++s)
                if (s >= LM - K) list_insert<LM>(l32, other[s]);
#pragma unroll
            for (int s = 0; s < LM; ++s) {
                if (s < LM - K) continue;
                const uint32_t k = l32[s];
                const uint64_t key = k == 0xFFFFFFFFu
                                         ? kNoKey
                                         : ((uint64_t)((k >> pb) - pk.off + na) << 32) |
                                               (uint32_t)borig[k & (uint32_t)(seg - 1)];
                if (first_seg)
                    l64[s] = key;
                else if (key < l64[LM - 1])
                    list_insert64<LM>(l64, key);
                l32[s] = 0xFFFFFFFFu;
            }
            first_seg = false;
        }
        __syncthreads();   // the tile (and, at a segment's end, borig) is consumed
        if (!next_same_seg) {
            // the next segment: of this leaf, or of the next probed leaf
            t0 = nt0;
            while (t0 >= cnt) {
                if (++pr >= nprobe) break;
                const int pl = leaf ^ pr;
                from = t.leaves[2 * pl];
                cnt = t.leaves[2 * pl + 1] - from;
                t0 = 0;
            }
            if (pr >= nprobe) break;
            load_segment();
            __syncthreads();
            np = min(kTP, cnt - t0);
            fetch_tile(np);
        }
    }
    if (!qv || h != 0) return;
    if (out.nn) {
#pragma unroll
        for (int s = 0; s < LM; ++s) {
            if (s < LM - K) continue;
            const int r = s - (LM - K);
            const uint64_t key = l64[s];
            const int jj = key == kNoKey ? -1 : (int)(uint32_t)key;
            if (out.knn) out.knn[(size_t)slot * K + r] = jj;
            if (r == 0) {
                out.nn[slot] = jj;
                if (out.dist) out.dist[slot] = key == kNoKey ? kFltMax : (float)__builtin_sqrt((double)(key >> 32));
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < LM; ++s)
            if (s >= LM - K) out.keys[(size_t)slot * K + (s - (LM - K))] = l64[s];
    }
}

// Block c's work record (chunks of kQ queries): the leaf l with
// chunk_off[l] <= c < chunk_off[l + 1]; blk_info[2c] = (leaf, first query
// slot in `order`, queries, first point of the leaf), blk_info[2c + 1].x =
// points of the leaf.
__global__ void k_block_info_q(const int *__restrict__ chunk_off, const int *__restrict__ offset,
                               const int *__restrict__ leaves, int nleaves, int nblk, int4 *__restrict__ blk_info) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nblk || c >= chunk_off[nleaves]) return;
    int lo = 0, hi = nleaves - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (chunk_off[m] <= c) lo = m; else hi = m - 1;
    }
    const int l = lo;
    const int qb = offset[l] + (c - chunk_off[l]) * kQ, from = leaves[2 * l];
    blk_info[2 * c] = make_int4(l, qb, min(kQ, offset[l + 1] - qb), from);
    blk_info[2 * c + 1] = make_int4(leaves[2 * l + 1] - from, 0, 0, 0);
}

// The two centred BF16 copies: out0[i] = element i, out1[i] = element i + 1
// (BF16 bits of byte - 128), zeros past the image; eight per thread.
__global__ void k_center_bf16(const uint8_t *__restrict__ img, long long n, long long npad,
                              uint16_t *__restrict__ out0, uint16_t *__restrict__ out1) {
    const long long i0 = 8 * ((long long)blockIdx.x * blockDim.x + threadIdx.x);
    if (i0 >= npad) return;
    auto cv = [](uint32_t x) -> uint32_t {   // exact: at most 8 significant bits
        return __float_as_uint((float)((int)x - 128)) >> 16;
    };
    if (i0 + 9 <= n && ((reinterpret_cast<uintptr_t>(img) + i0) & 7) == 0) {
        const uint2 w = *reinterpret_cast<const uint2 *>(img + i0);
        const uint32_t nx = img[i0 + 8];
        const uint32_t c0 = cv(w.x & 255u), c1 = cv((w.x >> 8) & 255u), c2 = cv((w.x >> 16) & 255u);
        const uint32_t c3 = cv(w.x >> 24), c4 = cv(w.y & 255u), c5 = cv((w.y >> 8) & 255u);
        const uint32_t c6 = cv((w.y >> 16) & 255u), c7 = cv(w.y >> 24), c8 = cv(nx);
        *reinterpret_cast<uint4 *>(out0 + i0) =
            make_uint4(c0 | (c1 << 16), c2 | (c3 << 16), c4 | (c5 << 16), c6 | (c7 << 16));
        *reinterpret_cast<uint4 *>(out1 + i0) =
            make_uint4(c1 | (c2 << 16), c3 | (c4 << 16), c5 | (c6 << 16), c7 | (c8 << 16));
    } else {
        for (long long i = i0; i < i0 + 8 && i < npad; ++i) {
            out0[i] = (uint16_t)(i < n ? cv(img[i]) : 0u);
            out1[i] = (uint16_t)(i + 1 < n ? cv(img[i + 1]) : 0u);
        }
    }
}

constexpr int kNormTile = 1024;   // row pixels per block of k_norm_rows

// hs[y * ncols + x] = sum over pixels x .. x + p - 1 of row y of the channel
// sums of (x - 128)^2: one block per (row, 1024-pixel tile), the per-pixel
// sums staged in LDS.
__global__ __launch_bounds__(256) void k_norm_rows(const uint8_t *__restrict__ img, int h, int w, int p,
                                                   uint32_t *__restrict__ hs) {
    __shared__ uint32_t sp[kNormTile + 16];
    const int ncols = w - p + 1;
    const int y = blockIdx.y, x0 = blockIdx.x * kNormTile;
    const int span = min(kNormTile + p - 1, w - x0);
    const uint8_t *row = img + ((long long)y * w + x0) * 3;
    for (int i = threadIdx.x; i < span; i += blockDim.x) {
        const int r = (int)row[3 * i] - 128, g = (int)row[3 * i + 1] - 128, b = (int)row[3 * i + 2] - 128;
        sp[i] = (uint32_t)(r * r + g * g + b * b);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNormTile && x0 + i < ncols; i += blockDim.x) {
        uint32_t s = 0;
        for (int k = 0; k < p; ++k) s += sp[i + k];
        hs[(long long)y * ncols + x0 + i] = s;
    }
}

// norm[y * ncols + x] = sum of hs over rows y .. y + p - 1
__global__ void k_norm_cols(const uint32_t *__restrict__ hs, int nrows, int ncols, int p, uint32_t *__restrict__ norm) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nrows * ncols) return;
    const int y = (int)(i / ncols), x = (int)(i - (long long)y * ncols);
    uint32_t s = 0;
    for (int r = 0; r < p; ++r) s += hs[(long long)(y + r) * ncols + x];
    norm[i] = s;
}

template <typename Fn>
hipError_t with_ks_lm(int ks, int lm_need, Fn &&fn) {
    auto pick_lm = [&](auto k) -> hipError_t {
        if (lm_need <= 5) return fn(k, std::integral_constant<int, 5>{});
        if (lm_need <= 8) return fn(k, std::integral_constant<int, 8>{});
        return fn(k, std::integral_constant<int, 32>{});
    };
    // the k-step counts of p = 5, 7, 8, 9, 10, 11, 12 exactly (other p run
    // the next one up, the steps past their own skipped)
    if (ks <= 5) return pick_lm(std::integral_constant<int, 5>{});
    if (ks <= 11) return pick_lm(std::integral_constant<int, 11>{});
    if (ks <= 12) return pick_lm(std::integral_constant<int, 12>{});
    if (ks <= 18) return pick_lm(std::integral_constant<int, 18>{});
    if (ks <= 20) return pick_lm(std::integral_constant<int, 20>{});
    if (ks <= 28) return pick_lm(std::integral_constant<int, 28>{});
    return pick_lm(std::integral_constant<int, 30>{});
}

PatchPack make_pack(int p) {
    PatchPack pk{};
    const int L = 3 * p, D = p * L;
    pk.p = p;
    pk.G = (L + 7) / 8;
    pk.ngroups = p * pk.G;
    pk.ks = (pk.ngroups + 1) / 2;
    pk.off = 16384u * (uint32_t)D;
    // position bits: ((65536 D) << pb) | (2^pb - 1) must stay below 0xFFFFFFFF
    const uint64_t vmax = 65536ull * (uint64_t)D;
    pk.pb = 8;
    while (pk.pb > 0 && ((vmax << pk.pb) | ((1ull << pk.pb) - 1)) >= 0xFFFFFFFFull) --pk.pb;
    return pk;
}

}  // namespace

int leafbf16_chunk() { return kQ; }

hipError_t leaf_blocks_q(const int *chunk_off, const int *offset, const int *leaves, int nleaves, int nblk,
                         int4 *blk_info, hipStream_t st) {
    hipLaunchKernelGGL(k_block_info_q, dim3((nblk + 255) / 256), dim3(256), 0, st, chunk_off, offset, leaves, nleaves,
                       nblk, blk_info);
    return hipGetLastError();
}

bool leafbf16_supported(int p) {
    if (p < 1 || p > 12) return false;   // images: fewer than 2^30 elements (checked by the caller)
    const PatchPack pk = make_pack(p);
    return pk.ks <= 30 && 2 * pk.ks <= kMaxGroups && pk.pb >= 5;
}

size_t leafbf16_scratch_bytes(int nq, int nleaves) {
    return ((size_t)nleaves + ((size_t)nq + kQ - 1) / kQ) * 2 * sizeof(int4);
}

// elements of one centred copy (the image and its zero padding); a buffer
// holds the two copies back to back
long long cimg_elems(int h, int w) { return 3LL * h * w + 16; }

hipError_t center_image(const uint8_t *img, int h, int w, uint16_t *out, hipStream_t st) {
    const long long n = 3LL * h * w, npad = cimg_elems(h, w);
    const long long nt = (npad + 7) / 8;
    hipLaunchKernelGGL(k_center_bf16, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, img, n, npad, out,
                       out + npad);
    return hipGetLastError();
}

hipError_t patch_norms_centered(const uint8_t *img, int h, int w, int p, uint32_t *tmp, uint32_t *norm,
                                hipStream_t st) {
    if (p > 16) return hipErrorInvalidValue;   // the LDS row tile's halo
    const int ncols = w - p + 1, nrows = h - p + 1;
    const dim3 g1((unsigned)((ncols + kNormTile - 1) / kNormTile), (unsigned)h);
    hipLaunchKernelGGL(k_norm_rows, g1, dim3(256), 0, st, img, h, w, p, tmp);
    const long long n2 = (long long)nrows * ncols;
    hipLaunchKernelGGL(k_norm_cols, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, tmp, nrows, ncols, p, norm);
    return hipGetLastError();
}

hipError_t leaf_bf16(const TreeView &t, const uint16_t *ca, int wa, const uint16_t *cb, int wb, int p,
                     const uint32_t *norm_a, const uint32_t *norm_b, const int *pts_order, const int *order,
                     const int *offset, const int *chunk_off, void *scratch, int nq, int nprobe, int K, int slot_div,
                     uint64_t *keys_out, int *knn, int *nn, float *dist, hipStream_t st, int ha, int hb, int a0) {
    if (nq == 0) return hipSuccess;
    if (!leafbf16_supported(p)) return hipErrorInvalidValue;
    const int nleaves = 1 << t.depth;
    if (nprobe < 1 || nprobe > nleaves || K < 1 || K > 32 || slot_div < 1) return hipErrorInvalidValue;
    const PatchPack pk = make_pack(p);
    const unsigned nblk = (unsigned)(nleaves + (nq + kQ - 1) / kQ);
    int4 *blk_info = static_cast<int4 *>(scratch);
    hipError_t e = leaf_blocks_q(chunk_off, offset, t.leaves, nleaves, (int)nblk, blk_info, st);
    if (e != hipSuccess) return e;
    const CImg ia{ca, (uint32_t)cimg_elems(ha, wa), wa, wa - p + 1}, ib{cb, (uint32_t)cimg_elems(hb, wb), wb, wb - p + 1};
    const Out out{knn, nn, dist, keys_out};
    return with_ks_lm(pk.ks, K, [&](auto ks, auto lm) -> hipError_t {
        constexpr int KSM = decltype(ks)::value;
        constexpr int LM = decltype(lm)::value;
        if constexpr (KSM == 5 || KSM == 11) {
            // p = 1-4 and 6 run these with the steps past their own skipped
            if (pk.ks != KSM) {
                hipLaunchKernelGGL((k_leaf_bf16<KSM, LM, false>), dim3(nblk), dim3(kThreads), 0, st, t, ia, ib, pk,
                                   norm_a, norm_b, pts_order, order, chunk_off, blk_info, nprobe, K, slot_div, a0, out);
                return hipGetLastError();
            }
        }
        if (pk.ks != KSM) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_leaf_bf16<KSM, LM, true>), dim3(nblk), dim3(kThreads), 0, st, t, ia, ib, pk, norm_a,
                           norm_b, pts_order, order, chunk_off, blk_info, nprobe, K, slot_div, a0, out);
        return hipGetLastError();
    });
}

}  // namespace annf
