// HIP kernels of the propagation-assisted k-d-tree NNF path, for gfx950.
//
// The semantics follow the reference's openmp core (annfmp/openmp/kdtree/
// src/*.c); the structure is GPU-first:
//   * the tree is built level-synchronously: per-node data box (min/max
//     reduction), first longest axis, radix select (11/11/10-bit LDS
//     histograms) of the rank-c/2 value, and an order-exact partition
//     computed from block scans (the closed form of helper.c:63-113, see
//     k_lvl_eqoff); the rows never move: a level moves the map from tree
//     positions to original rows and gathers rows through it; the last
//     levels run one subtree per block in LDS;
//   * the natural-leaf scan of every query is one pass over queries bucketed
//     by leaf, the leaf's points staged in LDS;
//   * row 0's exact search and the per-row propagation run one query per
//     wavefront: lanes score candidates, and the K best are extracted with
//     wave-wide minima of (distance, arrival) keys, which is exactly the
//     reference's sequential insert order (kdtree.c:502-532);
//   * the propagation is a persistent sweep: one wavefront owns one image
//     column and walks its rows; columns never wait on each other.
//
// Every FP32 distance is the sequential sum sum_j (p_j - q_j)^2 with one
// rounding per operation (the library is compiled with -ffp-contract=off),
// which is the reference's arithmetic (helper.c:164-179).
#include "annf_internal.h"
#include "annf_patch.h"

#include <type_traits>
#include <utility>

namespace annf {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// Order-preserving map float -> uint32 with -0 folded onto +0, so that keys
// compare like floats under '<' and '==' (finite values).
__device__ __forceinline__ uint32_t f2key(float f) {
    uint32_t u = __float_as_uint(f);
    if (u == 0x80000000u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ int readlane_i(int v, int lane) {
    return __builtin_amdgcn_readlane(v, lane);
}
__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    int lo = __shfl_xor((int)(uint32_t)v, m, kWave);
    int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, kWave);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

template <int DP>
__device__ __forceinline__ void load_vec(const float *__restrict__ p, float (&v)[DP]) {
#pragma unroll
    for (int c = 0; c < DP / 4; ++c) {
        float4 t = reinterpret_cast<const float4 *>(p)[c];
        v[4 * c + 0] = t.x;
        v[4 * c + 1] = t.y;
        v[4 * c + 2] = t.z;
        v[4 * c + 3] = t.w;
    }
}

// helper.c:164-179 order: t = p_j - q_j; s += t * t, j ascending.  The
// zero padding adds exact +0 terms after the last real dimension.
template <int DP>
__device__ __forceinline__ float dist_sq(const float *__restrict__ p, const float (&q)[DP]) {
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < DP / 4; ++c) {
        float4 v = reinterpret_cast<const float4 *>(p)[c];
        float t;
        t = v.x - q[4 * c + 0]; s = s + t * t;
        t = v.y - q[4 * c + 1]; s = s + t * t;
        t = v.z - q[4 * c + 2]; s = s + t * t;
        t = v.w - q[4 * c + 3]; s = s + t * t;
    }
    return s;
}

// Correctly rounded FP32 square root (the reference takes sqrt in double and
// stores a float, base.c:127 and base.c:371, which is the correctly rounded
// float root).  A first value comes from the FP64 root; it is then checked
// against the midpoints to its two neighbouring floats.  A midpoint has 25
// significant bits, so its square is exact in FP64, and no float equals the
// square of a midpoint, so there are no ties.
__device__ __forceinline__ float sqrt_rn(float x) {
    if (!(x > 0.0f) || !(x < INFINITY)) return (float)sqrt((double)x);
    float r = (float)sqrt((double)x);
    const double xd = (double)x;
    const float up = __uint_as_float(__float_as_uint(r) + 1u);
    const double mu = 0.5 * ((double)r + (double)up);
    if (mu * mu < xd) return up;
    const float dn = __uint_as_float(__float_as_uint(r) - 1u);
    const double md = 0.5 * ((double)r + (double)dn);
    if (md * md > xd) return dn;
    return r;
}

// [from, to) of leaf l computed from n and the depth alone: every node's
// segment splits at floor(c/2) (kdtree.c:105 and 114-115), so no table read
// (and no memory latency) is needed.
__device__ __forceinline__ void leaf_range(int l, int n, int depth, int &from, int &to) {
    int lo = 0, hi = n;
    for (int lev = depth - 1; lev >= 0; --lev) {
        const int mid = lo + ((hi - lo) >> 1);
        if ((l >> lev) & 1) lo = mid; else hi = mid;
    }
    from = lo;
    to = hi;
}

// Orders LDS accesses of the lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


template <typename Fn>
static hipError_t with_dp(int dpad, Fn &&fn) {
    switch (dpad) {
    case 4: return fn(std::integral_constant<int, 4>{});
    case 8: return fn(std::integral_constant<int, 8>{});
    case 12: return fn(std::integral_constant<int, 12>{});
    case 16: return fn(std::integral_constant<int, 16>{});
    case 20: return fn(std::integral_constant<int, 20>{});
    case 24: return fn(std::integral_constant<int, 24>{});
    case 28: return fn(std::integral_constant<int, 28>{});
    case 32: return fn(std::integral_constant<int, 32>{});
    case 48: return fn(std::integral_constant<int, 48>{});
    case 64: return fn(std::integral_constant<int, 64>{});
    default: return hipErrorInvalidValue;
    }
}

template <typename Fn>
static hipError_t with_dp_km(int dpad, int K, Fn &&fn) {
    return with_dp(dpad, [&](auto dp) {
        if (K <= 8) return fn(dp, std::integral_constant<int, 8>{});
        if (K <= 32) return fn(dp, std::integral_constant<int, 32>{});
        return fn(dp, std::integral_constant<int, 64>{});
    });
}

// padded dimension: a multiple of 4 up to 32, then 48 or 64 (zero padding
// adds exact +0 terms to every distance)
int dpad_of(int d) { return d <= 32 ? (d + 3) & ~3 : (d <= 48 ? 48 : 64); }
bool dpad_supported(int dpad) { return (dpad >= 4 && dpad <= 32 && dpad % 4 == 0) || dpad == 48 || dpad == 64; }
bool k_supported(int K) { return K >= 1 && K <= 64; }

static inline unsigned blocks_for(long long n, int tpb) {
    return (unsigned)((n + tpb - 1) / tpb);
}

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------

__global__ void k_pad(const float *__restrict__ src, int n, int d, int dpad, float *__restrict__ dst) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long tot = (long long)n * dpad;
    if (i >= tot) return;
    long long r = i / dpad;
    int c = (int)(i - r * dpad);
    dst[i] = c < d ? src[r * d + c] : 0.0f;
}

__global__ void k_iota(int *dst, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = i;
}

hipError_t pad_rows(const float *src, int n, int d, int dpad, float *dst, hipStream_t st) {
    long long tot = (long long)n * dpad;
    if (tot == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pad, dim3(blocks_for(tot, 256)), dim3(256), 0, st, src, n, d, dpad, dst);
    return hipGetLastError();
}

hipError_t iota(int *dst, int n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, st, dst, n);
    return hipGetLastError();
}

// Block-wide exclusive scan of an int2 over 256 threads (4 waves).
__device__ __forceinline__ int2 block_scan_int2(int2 v, int2 *wsum /* __shared__ [4] */) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int2 inc = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int tx = __shfl_up(inc.x, o, kWave);
        int ty = __shfl_up(inc.y, o, kWave);
        if (lane >= o) { inc.x += tx; inc.y += ty; }
    }
    if (lane == kWave - 1) wsum[wid] = inc;
    __syncthreads();
    int2 off = make_int2(0, 0);
    for (int w = 0; w < wid; ++w) { off.x += wsum[w].x; off.y += wsum[w].y; }
    __syncthreads();
    return make_int2(inc.x - v.x + off.x, inc.y - v.y + off.y);
}

// Block-wide exclusive scan of an int2 over 256 threads, also returning the
// block total.
__device__ __forceinline__ int2 block_scan_int2_tot(int2 v, int2 *wsum /* __shared__ [4] */, int2 &total) {
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int2 inc = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int tx = __shfl_up(inc.x, o, kWave);
        int ty = __shfl_up(inc.y, o, kWave);
        if (lane >= o) { inc.x += tx; inc.y += ty; }
    }
    if (lane == kWave - 1) wsum[wid] = inc;
    __syncthreads();
    int2 off = make_int2(0, 0), tot = make_int2(0, 0);
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
        if (w < wid) { off.x += wsum[w].x; off.y += wsum[w].y; }
        tot.x += wsum[w].x;
        tot.y += wsum[w].y;
    }
    __syncthreads();
    total = tot;
    return make_int2(inc.x - v.x + off.x, inc.y - v.y + off.y);
}

// ---------------------------------------------------------------------------
// tree build (level-synchronous)
// ---------------------------------------------------------------------------

constexpr int kBuildTPB = 256;
constexpr int kChunk = kBuildChunk;
constexpr int kBuildEPT = kChunk / kBuildTPB;
static_assert(kBuildEPT * kBuildTPB == kChunk, "a chunk is whole elements per thread");

struct LevelArgs {
    int nodes;        // 2^lvl
    int cpn;          // blocks (chunks of kChunk elements) per node
    const int *bnd;   // nodes + 1 segment bounds
    int node0;        // heap index of the level's first node
    int lvl;
    int rbox;         // copies of each node's box at this level (block chunk % rbox adds into one)
    int rbox_next;    // the same for the next level's boxes (filled by k_move)
};

// word of box copy `rep` of node `node`, dimension j, bound b (0 min, 1 max)
__device__ __forceinline__ size_t box_at(int node, int rbox, int rep, int dp, int j, int b) {
    return (((size_t)node * rbox + rep) * dp + j) * 2 + b;
}

__device__ __forceinline__ void chunk_range(const LevelArgs &L, int &node, int &lo, int &hi, int &s, int &e) {
    node = blockIdx.x / L.cpn;
    int chunk = blockIdx.x - node * L.cpn;
    lo = L.bnd[node];
    hi = L.bnd[node + 1];
    s = lo + chunk * kChunk;
    e = min(s + kChunk, hi);
}

// Radix select of the rank-c/2 key in three passes of 11, 11 and 10 bits.
// The first pass's node histogram has hist_rep(cpn) copies (its blocks add
// into copy chunk % R, which spreads the atomics of the wide top levels) and
// one block per node sums them and picks the digit; the later passes count
// only keys under the prefix picked so far, into one copy.
constexpr int kRBins = 2048;
constexpr int kHistRep = 8;
__host__ __device__ __forceinline__ int hist_rep(int cpn) { return cpn < kHistRep ? cpn : kHistRep; }
__host__ __device__ constexpr int pass_shift(int p) { return p == 0 ? 21 : (p == 1 ? 10 : 0); }
__host__ __device__ constexpr int pass_bins(int p) { return p == 2 ? 1024 : 2048; }

size_t level_hist_words(int nodes, int cpn) { return (size_t)nodes * hist_rep(cpn) * kRBins; }

size_t level_box_words(int nodes, int cpn, int dpad) { return (size_t)nodes * hist_rep(cpn) * dpad * 2; }

// Level set-up where the level kernels do not cover it: zero the level's
// first-pass histograms (the first level; later ones are cleared by the
// previous level) and reset the node boxes k_box accumulates into.
__global__ void k_lvl_prep(LevelArgs L, int dpad, uint32_t *box, uint32_t *H0, size_t hwords) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (box && i < (size_t)L.nodes * L.rbox * dpad) { box[2 * i] = 0xFFFFFFFFu; box[2 * i + 1] = 0u; }
    if (H0 && i < hwords) H0[i] = 0u;
}

// The digit of an NB-bin pass holding rank r of a node, and the number of
// keys in lower digits, from the R copies of its histogram at h (stride
// kRBins); every thread of the block gets the result.
template <int NB>
__device__ __forceinline__ void block_pick(const uint32_t *__restrict__ h, int R, uint32_t r, int2 *wsum,
                                           uint32_t *sres, uint32_t &digit, uint32_t &before) {
    constexpr int PER = NB / kBuildTPB;
    static_assert(PER % 4 == 0, "bins are read as uint4");
    const int tid = threadIdx.x;
    uint32_t c[PER], tot = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) c[j] = 0;
    // every copy's loads are issued before the first use
#pragma unroll
    for (int rr = 0; rr < kHistRep; ++rr) {
        if (rr < R) {
            const uint4 *p = reinterpret_cast<const uint4 *>(h + (size_t)rr * kRBins + tid * PER);
#pragma unroll
            for (int q = 0; q < PER / 4; ++q) {
                const uint4 v = p[q];
                c[4 * q + 0] += v.x;
                c[4 * q + 1] += v.y;
                c[4 * q + 2] += v.z;
                c[4 * q + 3] += v.w;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) tot += c[j];
    const uint32_t exc = (uint32_t)block_scan_int2(make_int2((int)tot, 0), wsum).x;
    if (r >= exc && r < exc + tot) {
        uint32_t b = exc;
        int dg = tid * PER + PER - 1;
        bool found = false;
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (!found && r < b + c[j]) {
                dg = tid * PER + j;
                found = true;
            }
            if (!found) b += c[j];
        }
        sres[0] = (uint32_t)dg;
        sres[1] = b;
    }
    __syncthreads();
    digit = sres[0];
    before = sres[1];
}

// Flushes a block's LDS histogram into copy `rep` of the node's histogram.
__device__ __forceinline__ void flush_hist(const uint32_t *h, uint32_t *g) {
    for (int i = threadIdx.x; i < kRBins; i += kBuildTPB) {
        const uint32_t c = h[i];
        if (c) atomicAdd(&g[i], c);
    }
}

// The pick of pass PASS for every node, one block per node: prefix and
// remaining rank (PASS 2: the split key vkey, into pre_out).
template <int PASS>
__global__ __launch_bounds__(kBuildTPB) void k_lvl_pick(LevelArgs L, const uint32_t *H, int R, const uint32_t *pre_in,
                                                        const int *rank_in, uint32_t *pre_out, int *rank_out) {
    __shared__ int2 wsum[kBuildTPB / kWave];
    __shared__ uint32_t sres[2];
    const int node = blockIdx.x;
    const uint32_t pre = PASS == 0 ? 0u : pre_in[node];
    const uint32_t r0 = PASS == 0 ? (uint32_t)((L.bnd[node + 1] - L.bnd[node]) / 2) : (uint32_t)rank_in[node];
    uint32_t digit, before;
    block_pick<pass_bins(PASS)>(H + (size_t)node * R * kRBins, R, r0, wsum, sres, digit, before);
    if (threadIdx.x == 0) {
        pre_out[node] = pre | (digit << pass_shift(PASS));
        if (rank_out) rank_out[node] = (int)(r0 - before);
    }
}

// Data box of each node (kdtree.c:379-394), as min/max of ordered keys.  The
// rows stay in their original order: tree position i is row orig[i] (i
// itself while orig is null).
template <int DP>
__global__ __launch_bounds__(kBuildTPB) void k_box(const float *__restrict__ pts, const int *__restrict__ orig,
                                                   LevelArgs L, int d, uint32_t *box) {
    int node, lo, hi, s, e;
    chunk_range(L, node, lo, hi, s, e);
    if (s >= e) return;
    __shared__ uint32_t smn[DP], smx[DP];
    if (threadIdx.x < DP) { smn[threadIdx.x] = 0xFFFFFFFFu; smx[threadIdx.x] = 0u; }
    __syncthreads();
    uint32_t mn[DP], mx[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) { mn[j] = 0xFFFFFFFFu; mx[j] = 0u; }
    for (int i = s + threadIdx.x; i < e; i += blockDim.x) {
        float v[DP];
        load_vec<DP>(pts + (size_t)(orig ? orig[i] : i) * DP, v);
#pragma unroll
        for (int j = 0; j < DP; ++j) {
            uint32_t k = f2key(v[j]);
            mn[j] = min(mn[j], k);
            mx[j] = max(mx[j], k);
        }
    }
#pragma unroll
    for (int j = 0; j < DP; ++j) {
#pragma unroll
        for (int m = kWave / 2; m >= 1; m >>= 1) {
            mn[j] = min(mn[j], (uint32_t)__shfl_xor((int)mn[j], m, kWave));
            mx[j] = max(mx[j], (uint32_t)__shfl_xor((int)mx[j], m, kWave));
        }
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int j = 0; j < DP; ++j) {
            atomicMin(&smn[j], mn[j]);
            atomicMax(&smx[j], mx[j]);
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < d) {
        const int rep = (blockIdx.x - node * L.cpn) % L.rbox;
        atomicMin(&box[box_at(node, L.rbox, rep, DP, threadIdx.x, 0)], smn[threadIdx.x]);
        atomicMax(&box[box_at(node, L.rbox, rep, DP, threadIdx.x, 1)], smx[threadIdx.x]);
    }
}

// Split axis (kdtree.c:369-418): cyclic = depth % d; longest = first axis of
// strictly largest FP32 width max - min, starting from 0.0.
__device__ __forceinline__ int node_axis(const LevelArgs &L, int node, int d, int dpad, const uint32_t *box,
                                         int splitting_type) {
    // called by a whole wavefront.  d <= 32: lane l takes dimension l % 32
    // and every other copy of the box, the halves combined by one shuffle;
    // d > 32: lane l takes dimension l and every copy
    if (splitting_type == 0) return L.lvl % d;
    const int lane = threadIdx.x & (kWave - 1);
    const bool halves = d <= 32;
    const int k = halves ? lane & 31 : lane;
    uint32_t kmn = 0xFFFFFFFFu, kmx = 0u;
    if (k < d)
        for (int rep = halves ? lane >> 5 : 0; rep < L.rbox; rep += halves ? 2 : 1) {
            kmn = min(kmn, box[box_at(node, L.rbox, rep, dpad, k, 0)]);
            kmx = max(kmx, box[box_at(node, L.rbox, rep, dpad, k, 1)]);
        }
    if (halves) {
        kmn = min(kmn, (uint32_t)__shfl_xor((int)kmn, 32, kWave));
        kmx = max(kmx, (uint32_t)__shfl_xor((int)kmx, 32, kWave));
    }
    // the first strictly largest FP32 width, starting from 0.0
    const float w = k < d ? key2f(kmx) - key2f(kmn) : 0.0f;
    float wmax = w;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, m, kWave));
    const unsigned long long hit = __ballot(k < d && (!halves || lane < 32) && w == wmax && w > 0.0f);
    return hit ? __ffsll(hit) - 1 : 0;
}

// First pass of a level: the node's split axis, the split-axis keys and the
// histogram of their top 11 bits.  Also clears this level's second- and
// third-pass histograms and resets the next level's boxes (nbox, filled by
// k_move) when given.
template <int DP>
__global__ __launch_bounds__(kBuildTPB) void k_lvl_hist0(const float *__restrict__ pts, const int *__restrict__ orig,
                                                         LevelArgs L, int d, int dpad,
                                                         const uint32_t *box, int splitting_type, int *nax,
                                                         int *axis_out, uint32_t *keys, uint32_t *H0, uint32_t *H1,
                                                         uint32_t *H2, size_t hwords, uint32_t *nbox) {
    __shared__ uint32_t h[kRBins];
    __shared__ int sax;
    {
        const size_t gt = (size_t)blockIdx.x * kBuildTPB + threadIdx.x, gn = (size_t)gridDim.x * kBuildTPB;
        for (size_t i = gt; i < (size_t)L.nodes * kRBins; i += gn) {
            H1[i] = 0u;
            H2[i] = 0u;
        }
        if (nbox)
            for (size_t i = gt; i < (size_t)2 * L.nodes * L.rbox_next * dpad; i += gn) {
                nbox[2 * i] = 0xFFFFFFFFu;
                nbox[2 * i + 1] = 0u;
            }
    }
    int node, lo, hi, s, e;
    chunk_range(L, node, lo, hi, s, e);
    if (s >= e) return;
    for (int i = threadIdx.x; i < kRBins; i += kBuildTPB) h[i] = 0u;
    if (threadIdx.x < kWave) {
        const int ax = node_axis(L, node, d, dpad, box, splitting_type);
        if (threadIdx.x == 0) {
            sax = ax;
            if (s == lo) {
                nax[node] = ax;
                axis_out[L.node0 + node] = ax;
            }
        }
    }
    __syncthreads();
    const int ax = sax;
    // tree position i is row orig[i]: every row index of the thread is read
    // before the gathers, every gather before the first key
    int row[kBuildEPT];
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) {
        const int i = s + (int)threadIdx.x + u * kBuildTPB;
        row[u] = i < e ? (orig ? orig[i] : i) : -1;
    }
    float v[kBuildEPT];
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) v[u] = row[u] >= 0 ? pts[(size_t)row[u] * DP + ax] : 0.0f;
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) {
        if (row[u] < 0) continue;
        const uint32_t k = f2key(v[u]);
        keys[s + (int)threadIdx.x + u * kBuildTPB] = k;
        atomicAdd(&h[k >> pass_shift(0)], 1u);
    }
    __syncthreads();
    const int R = hist_rep(L.cpn), rep = (blockIdx.x - node * L.cpn) % R;
    flush_hist(h, H0 + ((size_t)node * R + rep) * kRBins);
}

// Pass PASS (1 or 2): the histogram of the next digit over the keys that
// match the prefix picked so far, into one copy per node (few keys match, so
// the atomics do not contend).  Pass 2 first makes pass 1's pick itself, every
// block from the complete histogram (the same result in every block; the
// node's first block stores it).
template <int PASS>
__global__ __launch_bounds__(kBuildTPB) void k_lvl_pass(LevelArgs L, const uint32_t *keys, const uint32_t *Hp,
                                                        uint32_t *Hc, const uint32_t *pre_in, const int *rank_in,
                                                        uint32_t *pre_out, int *rank_out) {
    __shared__ uint32_t h[kRBins];
    __shared__ int2 wsum[kBuildTPB / kWave];
    __shared__ uint32_t sres[2];
    int node, lo, hi, s, e;
    chunk_range(L, node, lo, hi, s, e);
    if (s >= e) return;
    for (int i = threadIdx.x; i < kRBins; i += kBuildTPB) h[i] = 0u;
    uint32_t pre = pre_in[node];
    if (PASS == 2) {
        const uint32_t r0 = (uint32_t)rank_in[node];
        uint32_t digit, before;
        block_pick<pass_bins(1)>(Hp + (size_t)node * kRBins, 1, r0, wsum, sres, digit, before);
        pre |= digit << pass_shift(1);
        if (s == lo && threadIdx.x == 0) {
            pre_out[node] = pre;
            rank_out[node] = (int)(r0 - before);
        }
    }
    __syncthreads();
    for (int i = s + threadIdx.x; i < e; i += kBuildTPB) {
        const uint32_t k = keys[i];
        if (((k ^ pre) >> pass_shift(PASS - 1)) == 0u)
            atomicAdd(&h[(k >> pass_shift(PASS)) & (uint32_t)(pass_bins(PASS) - 1)], 1u);
    }
    __syncthreads();
    flush_hist(h, Hc + (size_t)node * kRBins);
}

// The last pick (the node's split key, made by every block from the complete
// pass-2 histogram; the node's first block stores it in vkey) and per-chunk
// counts of keys below / equal to it.
__global__ __launch_bounds__(kBuildTPB) void k_lvl_count(LevelArgs L, const uint32_t *keys, const uint32_t *H2,
                                                         const uint32_t *pre_in, const int *rank_in, uint32_t *vkey,
                                                         int2 *cnt) {
    __shared__ int2 wsum[kBuildTPB / kWave];
    __shared__ uint32_t sres[2];
    int node, lo, hi, s, e;
    chunk_range(L, node, lo, hi, s, e);
    if (s >= e) {
        if (threadIdx.x == 0) cnt[blockIdx.x] = make_int2(0, 0);
        return;
    }
    uint32_t digit, before;
    block_pick<pass_bins(2)>(H2 + (size_t)node * kRBins, 1, (uint32_t)rank_in[node], wsum, sres, digit, before);
    const uint32_t vk = pre_in[node] | digit;
    if (s == lo && threadIdx.x == 0) vkey[node] = vk;
    int lt = 0, eq = 0;
    for (int i = s + threadIdx.x; i < e; i += kBuildTPB) {
        const uint32_t k = keys[i];
        lt += k < vk;
        eq += k == vk;
    }
    const int2 ex = block_scan_int2(make_int2(lt, eq), wsum);
    if (threadIdx.x == kBuildTPB - 1) cnt[blockIdx.x] = make_int2(ex.x + lt, ex.y + eq);
}

// Order-exact partition of one node segment (helper.c:63-113), in closed
// form.  With L = #(x < v), E = #(x == v), g = c - L:
//   x <  v : slot = rank among '<' elements (segment order)
//   x == v : slot = L + t, t = (E-1) - rank among equal elements
//   x >  v : G-offset go = (g-1) - rank among '>=' elements; follow
//            go -> eqoff[go] while go < E; slot = L + go
// where eqoff[t] is the G-offset of the t-th equal element in G order.
// This kernel sums the chunk counts into the block's offsets (cnt_off) and
// the node totals (tot) and records eqoff; k_move moves the records.  It also
// clears the next level's first-pass histograms (H0n, hwords_n words).
template <int DP>
__global__ __launch_bounds__(kBuildTPB) void k_lvl_eqoff(LevelArgs L, const uint32_t *keys, const uint32_t *vkey,
                                                         const int2 *cnt, int2 *cnt_off, int2 *tot, int *eqoff,
                                                         uint32_t *H0n, size_t hwords_n, int *err) {
    __shared__ int2 wsum[kBuildTPB / kWave];
    {
        const size_t gt = (size_t)blockIdx.x * kBuildTPB + threadIdx.x, gn = (size_t)gridDim.x * kBuildTPB;
        for (size_t i = gt; i < hwords_n; i += gn) H0n[i] = 0u;
    }
    int node, lo, hi, s, e;
    chunk_range(L, node, lo, hi, s, e);
    if (s >= e) return;
    const int ch = blockIdx.x - node * L.cpn;
    int2 before = make_int2(0, 0), all = make_int2(0, 0);
    for (int j = threadIdx.x; j < L.cpn; j += kBuildTPB) {
        const int2 v = cnt[(size_t)node * L.cpn + j];
        all.x += v.x;
        all.y += v.y;
        if (j < ch) { before.x += v.x; before.y += v.y; }
    }
    int2 sum_b, sum_a;
    block_scan_int2_tot(before, wsum, sum_b);
    block_scan_int2_tot(all, wsum, sum_a);
    if (threadIdx.x == 0) {
        cnt_off[blockIdx.x] = sum_b;
        if (ch == 0) tot[node] = sum_a;
    }
    const int c = hi - lo;
    const int Lt = sum_a.x, E = sum_a.y, g = c - Lt;
    const uint32_t vk = vkey[node];
    const int i0 = s + threadIdx.x * kBuildEPT;
    uint32_t k[kBuildEPT];
    int nlt = 0, neq = 0;
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) {
        const int i = i0 + u;
        k[u] = (i < e) ? keys[i] : 0u;
        if (i < e) { nlt += k[u] < vk; neq += k[u] == vk; }
    }
    const int2 ex = block_scan_int2(make_int2(nlt, neq), wsum);
    int r_lt = sum_b.x + ex.x;
    int r_eq = sum_b.y + ex.y;
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) {
        const int i = i0 + u;
        if (i >= e) break;
        if (k[u] < vk) {
            ++r_lt;
        } else if (k[u] == vk) {
            const int go = (g - 1) - (i - lo - r_lt);
            const int t = (E - 1) - r_eq++;
            if (t < 0 || t >= c) { atomicOr(err, 2); continue; }
            eqoff[lo + t] = go;
        }
    }
}

// The moving pass of the partition (the slots of k_lvl_eqoff).  The rows never
// move: every level reads them through orig (tree position -> original
// row), so this pass moves orig alone.  The block first computes every
// point's slot (eight consecutive points per thread, so the segment-order
// ranks come from one block scan) and writes orig in the new order; with
// nbox (DP/4 a power of two) it then gathers the rows, one float4 per lane,
// and accumulates the two children's data boxes (the next level's k_box).
template <int DP>
__global__ __launch_bounds__(kBuildTPB) void k_move(const float *__restrict__ pts, const int *__restrict__ orig_in,
                                                    LevelArgs L, const uint32_t *keys, const uint32_t *vkey,
                                                    const int2 *cnt_off, const int2 *tot, const int *eqoff,
                                                    int *__restrict__ orig_out, const int *nax, float *split,
                                                    int *err, uint32_t *nbox, int d) {
    constexpr int P4 = DP / 4;
    constexpr bool kBoxOk = (P4 & (P4 - 1)) == 0;
    __shared__ int2 wsum[kBuildTPB / kWave];
    __shared__ int slots[kChunk];
    __shared__ int srow[kChunk];
    __shared__ uint32_t sbox[2][DP][2];
    int node, lo, hi, s, e;
    chunk_range(L, node, lo, hi, s, e);
    if (s >= e) return;
    const int c = hi - lo, half = c / 2;
    const int Lt = tot[node].x, E = tot[node].y, g = c - Lt;
    const uint32_t vk = vkey[node];
    const int ax = nax[node];
    const bool boxes = kBoxOk && nbox != nullptr;
    if (boxes && threadIdx.x < 2 * DP) {
        sbox[threadIdx.x / DP][threadIdx.x % DP][0] = 0xFFFFFFFFu;
        sbox[threadIdx.x / DP][threadIdx.x % DP][1] = 0u;
    }
    // thread t takes the kBuildEPT consecutive points s + t * kBuildEPT, ...:
    // one block scan of the per-thread counts gives every rank
    const int i0 = s + (int)threadIdx.x * kBuildEPT;
    uint32_t kk[kBuildEPT];
    int nlt = 0, neq = 0;
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) {
        const int i = i0 + u;
        kk[u] = i < e ? keys[i] : 0u;
        nlt += i < e && kk[u] < vk;
        neq += i < e && kk[u] == vk;
    }
    int2 sum;
    const int2 ex = block_scan_int2_tot(make_int2(nlt, neq), wsum, sum);
    const int2 run = cnt_off[blockIdx.x];
    int r_lt = run.x + ex.x, r_eq = run.y + ex.y;   // points below / equal to the split key before point i
#pragma unroll
    for (int u = 0; u < kBuildEPT; ++u) {
        const int i = i0 + u;
        const uint32_t k = kk[u];
        int slot = -1;
        if (i < e) {
            const int off = i - lo;
            if (k < vk) {
                slot = r_lt++;
            } else if (k == vk) {
                slot = Lt + (E - 1) - r_eq++;
            } else {
                int p = (g - 1) - (off - r_lt);
                // the chain strictly increases, so it ends within E steps
                for (int it = 0; p < E && it <= E; ++it) p = eqoff[lo + p];
                slot = Lt + p;
            }
            if (slot < 0 || slot >= c) {   // cannot happen for a correct partition
                atomicOr(err, 1);
                slot = -1;
            } else {
                const int row = orig_in ? orig_in[i] : i;
                orig_out[lo + slot] = row;
                srow[i - s] = row;
                if (slot == half) split[L.node0 + node] = pts[(size_t)row * DP + ax];
            }
        }
        slots[i - s] = slot;
    }
    if (!boxes) return;
    __syncthreads();
    const float4 *src4 = reinterpret_cast<const float4 *>(pts);
    uint32_t mn0[4], mx0[4], mn1[4], mx1[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { mn0[q] = mn1[q] = 0xFFFFFFFFu; mx0[q] = mx1[q] = 0u; }
    const int nf = (e - s) * P4;
    constexpr int kUn = 8;   // gathers in flight per thread
    for (int f0 = threadIdx.x; f0 < nf; f0 += kUn * kBuildTPB) {
        int sl[kUn];
        float4 v[kUn];
#pragma unroll
        for (int u = 0; u < kUn; ++u) {
            const int f = f0 + u * kBuildTPB;
            const int el = f / P4, part = f - el * P4;
            sl[u] = f < nf ? slots[el] : -1;
            if (sl[u] >= 0) v[u] = src4[(size_t)srow[el] * P4 + part];
        }
#pragma unroll
        for (int u = 0; u < kUn; ++u) {
            if (sl[u] < 0) continue;
            const uint32_t k0 = f2key(v[u].x), k1 = f2key(v[u].y), k2 = f2key(v[u].z), k3 = f2key(v[u].w);
            if (sl[u] < half) {
                mn0[0] = min(mn0[0], k0); mx0[0] = max(mx0[0], k0);
                mn0[1] = min(mn0[1], k1); mx0[1] = max(mx0[1], k1);
                mn0[2] = min(mn0[2], k2); mx0[2] = max(mx0[2], k2);
                mn0[3] = min(mn0[3], k3); mx0[3] = max(mx0[3], k3);
            } else {
                mn1[0] = min(mn1[0], k0); mx1[0] = max(mx1[0], k0);
                mn1[1] = min(mn1[1], k1); mx1[1] = max(mx1[1], k1);
                mn1[2] = min(mn1[2], k2); mx1[2] = max(mx1[2], k2);
                mn1[3] = min(mn1[3], k3); mx1[3] = max(mx1[3], k3);
            }
        }
    }
    // lanes with the same part (lane % P4) hold the same four dimensions
    // (kUn * kBuildTPB is a multiple of P4, so a lane's part never changes)
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int m = P4; m < kWave; m <<= 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            mn0[q] = min(mn0[q], (uint32_t)__shfl_xor((int)mn0[q], m, kWave));
            mx0[q] = max(mx0[q], (uint32_t)__shfl_xor((int)mx0[q], m, kWave));
            mn1[q] = min(mn1[q], (uint32_t)__shfl_xor((int)mn1[q], m, kWave));
            mx1[q] = max(mx1[q], (uint32_t)__shfl_xor((int)mx1[q], m, kWave));
        }
    }
    if (lane < P4) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = 4 * lane + q;
            atomicMin(&sbox[0][j][0], mn0[q]);
            atomicMax(&sbox[0][j][1], mx0[q]);
            atomicMin(&sbox[1][j][0], mn1[q]);
            atomicMax(&sbox[1][j][1], mx1[q]);
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * d) {
        const int ch = threadIdx.x / d, j = threadIdx.x - ch * d;
        const size_t o = box_at(2 * node + ch, L.rbox_next, blockIdx.x % L.rbox_next, DP, j, 0);
        atomicMin(&nbox[o], sbox[ch][j][0]);
        atomicMax(&nbox[o + 1], sbox[ch][j][1]);   // copy blockIdx.x % rbox_next
    }
}

// Tree-order rows when no subtree pass writes them: out[i] = pts[orig[i]].
template <int DP>
__global__ void k_gather_rows(const float *__restrict__ pts, const int *__restrict__ orig, int n,
                              float *__restrict__ out) {
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= (long long)n * (DP / 4)) return;
    const long long i = f / (DP / 4);
    const int part = (int)(f - i * (DP / 4));
    reinterpret_cast<float4 *>(out)[f] = reinterpret_cast<const float4 *>(pts)[(size_t)orig[i] * (DP / 4) + part];
}

// ---------------------------------------------------------------------------
// the last levels of the build, one subtree per block
// ---------------------------------------------------------------------------
// Once every node of a level holds at most sub_max<DP> points, one
// 1024-thread block builds the remaining (at most kSubLevels) levels of each
// node's subtree with the same rules as the level kernels above -- data box
// of ordered keys, first strictly longest FP32 width (or level % d), radix
// select of the rank-c/2 key in four 8-bit passes, the closed-form partition
// order -- with the node's points resident in LDS (read once) and a 16-bit
// index permutation rearranged per level; the points are written once, in
// final order, at the end.
constexpr int kSubLevels = 4;
constexpr int kSubTPB = 1024;
constexpr int kSubNodes = 1 << (kSubLevels - 1);   // sub-nodes of the last level done in a block
constexpr int kSubMaxAll = 2048;
template <int DP>
constexpr int sub_max() { return DP <= 16 ? 2048 : (DP <= 32 ? 1024 : 512); }   // at most 128 KiB of points
static_assert(kSubNodes * 256 == kSubMaxAll, "radix histograms alias the eqoff array");

int subtree_max(int dpad) { return dpad <= 16 ? 2048 : (dpad <= 32 ? 1024 : 512); }

__device__ __forceinline__ int sub_of(const int *sb, int nsub, int i) {
    int a = 0, b = nsub - 1;
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (sb[m] <= i) a = m; else b = m - 1;
    }
    return a;
}

template <int DP>
__global__ __launch_bounds__(kSubTPB) void k_subtree(const float *__restrict__ pts_in, const int *__restrict__ orig_in,
                                                     const int *__restrict__ src_idx, const int *__restrict__ bnd0,
                                                     int L0, int R, int d, int splitting_type, float *__restrict__ pts_out,
                                                     int *__restrict__ orig_out, int *__restrict__ axis_out,
                                                     float *__restrict__ split_out, int *err,
                                                     int *__restrict__ leaves, int *__restrict__ pli) {
    constexpr int SM = sub_max<DP>();
    constexpr int EPT = (SM + kSubTPB - 1) / kSubTPB;   // points per thread in the partition scan
    // the node's points, level-L0 order; dimension j of point p is stored at
    // p * DP + (j + p) % DP, so the lanes reading one dimension of different
    // points spread over the LDS banks
    __shared__ float dataf[SM * DP];
    __shared__ uint16_t perm[2][SM];
    __shared__ uint32_t keys[SM];
    __shared__ int hist_eq[kSubMaxAll];                 // radix histograms, then eqoff
    __shared__ uint32_t sbox[kSubNodes][DP][2];
    __shared__ int sb[kSubNodes + 1];                   // sub-node bounds (local offsets)
    __shared__ uint8_t sid[SM];                         // sub-node of each local position
    __shared__ int sax[kSubNodes];
    __shared__ uint32_t spre[kSubNodes];
    __shared__ int srank[kSubNodes];
    __shared__ int PL[kSubNodes + 1], PE[kSubNodes + 1];
    __shared__ int wlt[kSubTPB / kWave], weq[kSubTPB / kWave];
    auto at = [](int p, int j) { return p * DP + (j + p) % DP; };

    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int node = blockIdx.x;
    const int lo = bnd0[node], c0 = bnd0[node + 1] - lo;
    {
        // point i of the node is row src_idx[lo + i] (or lo + i)
        for (int e = tid; e < c0 * (DP / 4); e += kSubTPB) {
            const int i = e / (DP / 4), j = 4 * (e - i * (DP / 4));
            const size_t g = src_idx ? (size_t)src_idx[lo + i] : (size_t)(lo + i);
            const float4 v = reinterpret_cast<const float4 *>(pts_in + g * DP)[j / 4];
            dataf[at(i, j + 0)] = v.x;
            dataf[at(i, j + 1)] = v.y;
            dataf[at(i, j + 2)] = v.z;
            dataf[at(i, j + 3)] = v.w;
        }
    }
    for (int i = tid; i < c0; i += kSubTPB) perm[0][i] = (uint16_t)i;
    if (tid == 0) {
        sb[0] = 0;
        sb[1] = c0;
    }
    int cur = 0;
    __syncthreads();
    for (int r = 0; r < R; ++r) {
        const int nsub = 1 << r;
        const int lvl = L0 + r;
        const int heap0 = (1 << lvl) - 1 + node * nsub;
        if (tid < nsub * DP) {
            sbox[tid / DP][tid % DP][0] = 0xFFFFFFFFu;
            sbox[tid / DP][tid % DP][1] = 0u;
        }
        for (int i = tid; i < c0; i += kSubTPB) sid[i] = (uint8_t)sub_of(sb, nsub, i);
        __syncthreads();
        // (A) data boxes: every (sub-node, dimension) pair gets T threads
        // (a power of two, T | 64) that scan its points; no atomics
        if (splitting_type != 0) {
            const int P = nsub * d;
            int T = 64;
            while (T > 1 && T * P > kSubTPB) T >>= 1;
            const int pair = tid / T, lip = tid % T;
            uint32_t mn = 0xFFFFFFFFu, mx = 0u;
            int s2 = 0, j = 0;
            if (pair < P) {
                s2 = pair / d;
                j = pair - s2 * d;
                for (int i = sb[s2] + lip; i < sb[s2 + 1]; i += T) {
                    const uint32_t k = f2key(dataf[at(perm[cur][i], j)]);
                    mn = min(mn, k);
                    mx = max(mx, k);
                }
            }
            for (int m = T >> 1; m >= 1; m >>= 1) {
                mn = min(mn, (uint32_t)__shfl_xor((int)mn, m, kWave));
                mx = max(mx, (uint32_t)__shfl_xor((int)mx, m, kWave));
            }
            if (pair < P && lip == 0) {
                sbox[s2][j][0] = mn;
                sbox[s2][j][1] = mx;
            }
        }
        __syncthreads();
        // (B) split axis per sub-node
        if (tid < nsub) {
            int ax = 0;
            if (splitting_type == 0) {
                ax = lvl % d;
            } else {
                float longest = 0.0f;
                for (int k = 0; k < d; ++k) {
                    const float w = key2f(sbox[tid][k][1]) - key2f(sbox[tid][k][0]);
                    if (w > longest) { longest = w; ax = k; }
                }
            }
            sax[tid] = ax;
            axis_out[heap0 + tid] = ax;
            spre[tid] = 0u;
            srank[tid] = (sb[tid + 1] - sb[tid]) / 2;
        }
        __syncthreads();
        // (C) split-axis keys
        for (int i = tid; i < c0; i += kSubTPB)
            keys[i] = f2key(dataf[at(perm[cur][i], sax[sid[i]])]);
        // (D) radix select of the rank-c/2 key of every sub-node
        for (int pass = 0; pass < 4; ++pass) {
            for (int i = tid; i < nsub * 256; i += kSubTPB) hist_eq[i] = 0;
            __syncthreads();
            const int dsh = 24 - 8 * pass;
            for (int i = tid; i < c0; i += kSubTPB) {
                const int s2 = sid[i];
                const uint32_t k = keys[i];
                if (pass == 0 || ((k ^ spre[s2]) >> (dsh + 8)) == 0u)
                    atomicAdd(&hist_eq[s2 * 256 + ((k >> dsh) & 255u)], 1);
            }
            __syncthreads();
            if (wid < nsub) {
                const int *h = hist_eq + wid * 256;
                const uint32_t c_0 = h[lane * 4 + 0], c_1 = h[lane * 4 + 1], c_2 = h[lane * 4 + 2],
                               c_3 = h[lane * 4 + 3];
                const uint32_t tot = c_0 + c_1 + c_2 + c_3;
                uint32_t inc = tot;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t t = (uint32_t)__shfl_up((int)inc, o, kWave);
                    if (lane >= o) inc += t;
                }
                const uint32_t exc = inc - tot;
                const uint32_t rr = (uint32_t)srank[wid];
                const bool mine = rr >= exc && rr < inc;
                int digit = 0;
                uint32_t before = exc;
                if (mine) {
                    if (rr < exc + c_0) { digit = lane * 4 + 0; before = exc; }
                    else if (rr < exc + c_0 + c_1) { digit = lane * 4 + 1; before = exc + c_0; }
                    else if (rr < exc + c_0 + c_1 + c_2) { digit = lane * 4 + 2; before = exc + c_0 + c_1; }
                    else { digit = lane * 4 + 3; before = exc + c_0 + c_1 + c_2; }
                }
                const unsigned long long bal = __ballot(mine);
                const int owner = __ffsll(bal) - 1;
                digit = readlane_i(digit, owner);
                before = (uint32_t)readlane_i((int)before, owner);
                if (lane == 0) {
                    spre[wid] |= (uint32_t)digit << dsh;
                    srank[wid] = (int)(rr - before);
                }
            }
            __syncthreads();
        }
        // (E) '<' and '==' counts in segment order: EPT contiguous points per
        // thread, block-wide exclusive scan over the concatenated segments
        int fl[EPT], fe[EPT];
        int tl = 0, te = 0;
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int i = EPT * tid + u;
            fl[u] = fe[u] = 0;
            if (i < c0) {
                const uint32_t v = spre[sid[i]];
                fl[u] = keys[i] < v;
                fe[u] = keys[i] == v;
            }
            tl += fl[u];
            te += fe[u];
        }
        int il = tl, ie = te;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int a = __shfl_up(il, o, kWave), b2 = __shfl_up(ie, o, kWave);
            if (lane >= o) { il += a; ie += b2; }
        }
        if (lane == kWave - 1) { wlt[wid] = il; weq[wid] = ie; }
        __syncthreads();
        int ol = 0, oe = 0;
        for (int w = 0; w < wid; ++w) { ol += wlt[w]; oe += weq[w]; }
        int xl = il - tl + ol, xe = ie - te + oe;   // exclusive prefix of point EPT*tid
        int pl[EPT], pe[EPT];
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            pl[u] = xl;
            pe[u] = xe;
            const int i = EPT * tid + u;
            for (int s2 = 0; s2 <= nsub; ++s2)
                if (sb[s2] == i && i < c0) { PL[s2] = xl; PE[s2] = xe; }
            xl += fl[u];
            xe += fe[u];
        }
        if (tid == kSubTPB - 1) { PL[nsub] = xl; PE[nsub] = xe; }
        __syncthreads();
        // (F) eqoff[t] = G-offset of the t-th equal point in G order
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int i = EPT * tid + u;
            if (i >= c0 || !fe[u]) continue;
            const int s2 = sid[i];
            const int c = sb[s2 + 1] - sb[s2], off = i - sb[s2];
            const int Ls = PL[s2 + 1] - PL[s2], Es = PE[s2 + 1] - PE[s2];
            const int rl = pl[u] - PL[s2], re = pe[u] - PE[s2];
            const int go = (c - Ls - 1) - (off - rl);
            const int t = (Es - 1) - re;
            if (t < 0 || t >= c) { atomicOr(err, 2); continue; }
            hist_eq[sb[s2] + t] = go;
        }
        __syncthreads();
        // (G) slots
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
            const int i = EPT * tid + u;
            if (i >= c0) continue;
            const int s2 = sid[i];
            const int c = sb[s2 + 1] - sb[s2], off = i - sb[s2];
            const int Ls = PL[s2 + 1] - PL[s2], Es = PE[s2 + 1] - PE[s2];
            const int rl = pl[u] - PL[s2], re = pe[u] - PE[s2];
            int slot;
            if (fl[u]) {
                slot = rl;
            } else {
                const int go = (c - Ls - 1) - (off - rl);
                if (fe[u]) {
                    slot = Ls + (Es - 1) - re;
                } else {
                    int q = go;
                    for (int it = 0; q < Es && it <= Es; ++it) q = hist_eq[sb[s2] + q];
                    slot = Ls + q;
                }
            }
            if (slot < 0 || slot >= c) { atomicOr(err, 1); continue; }
            const int pidx = perm[cur][i];
            perm[cur ^ 1][sb[s2] + slot] = (uint16_t)pidx;
            if (slot == c / 2) split_out[heap0 + s2] = dataf[at(pidx, sax[s2])];
        }
        __syncthreads();
        // children bounds for the next level: the left child gets c/2 points
        // (kdtree.c:114-115); 2 * nsub <= kSubNodes there
        if (r + 1 < R) {
            int b_lo = 0, b_hi = 0;
            if (tid < nsub) {
                b_lo = sb[tid];
                b_hi = sb[tid + 1];
            }
            __syncthreads();
            if (tid < nsub) {
                sb[2 * tid] = b_lo;
                sb[2 * tid + 1] = b_lo + (b_hi - b_lo) / 2;
            }
            if (tid == 0) sb[2 * nsub] = c0;
        }
        cur ^= 1;
        __syncthreads();
    }
    // the points in final order (pts_out may be null: the full-patch path
    // never reads them)
    float4 *dst = pts_out ? reinterpret_cast<float4 *>(pts_out + (size_t)lo * DP) : nullptr;
    for (int e = tid; dst && e < c0 * (DP / 4); e += kSubTPB) {
        const int i = e / (DP / 4), j = 4 * (e - i * (DP / 4));
        const int pi = perm[cur][i];
        dst[e] = make_float4(dataf[at(pi, j + 0)], dataf[at(pi, j + 1)], dataf[at(pi, j + 2)], dataf[at(pi, j + 3)]);
    }
    // final order, and (when asked) the leaf bounds and point -> leaf map:
    // sub-node s2 of the last level done splits at c/2 into leaves
    // (node << R) + 2 s2 and + 1 (kdtree.c:86-100, 114-115)
    const int nlast = 1 << (R - 1);
    for (int i = tid; i < c0; i += kSubTPB) {
        const int src = lo + perm[cur][i];
        const int o = orig_in ? orig_in[src] : src;
        orig_out[lo + i] = o;
        if (pli) {
            const int s2 = sid[i];
            const int half = (sb[s2 + 1] - sb[s2]) / 2;
            pli[o] = (node << R) + 2 * s2 + (i - sb[s2] >= half ? 1 : 0);
        }
    }
    if (leaves && tid < nlast) {
        const int l0 = (node << R) + 2 * tid, mid = sb[tid] + (sb[tid + 1] - sb[tid]) / 2;
        leaves[2 * l0] = lo + sb[tid];
        leaves[2 * l0 + 1] = lo + mid;
        leaves[2 * l0 + 2] = lo + mid;
        leaves[2 * l0 + 3] = lo + sb[tid + 1];
    }
}

// Leaf bounds and point -> leaf map (kdtree.c:86-100).
__global__ void k_leaves(const int *bnd, int nleaves, const int *orig, int n, int *leaves, int *pli) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nleaves) { leaves[2 * i] = bnd[i]; leaves[2 * i + 1] = bnd[i + 1]; }
    if (i < n) {
        int a = 0, b = nleaves - 1;
        while (a < b) {
            int m = (a + b + 1) >> 1;
            if (bnd[m] <= i) a = m; else b = m - 1;
        }
        pli[orig[i]] = a;
    }
}

hipError_t build_tree(int n, int d, int dpad, int depth, int splitting_type,
                      const int *host_bnd, const int *host_bnd_off, const int *host_cpn,
                      const int *host_maxc, BuildScratch &s, int *axis, float *split, int *leaves, int *pli,
                      TreeView &t, hipStream_t st, bool need_pts) {
    (void)host_bnd;
    // the n x dpad rows stay in s.pts_a in their original order; the levels
    // move orig (tree position -> row) and read the rows through it; the
    // tree-order rows are written once at the end, into s.pts_b
    const float *rows = s.pts_a;
    float *rows_out = s.pts_b;
    int *oin = s.orig_a, *oout = s.orig_b;
    hipError_t err = hipSuccess;
    // first level whose nodes all fit one subtree block, at most kSubLevels
    // above the leaves
    int L0 = depth;
    for (int lvl = std::max(0, depth - kSubLevels); lvl < depth; ++lvl)
        if (host_maxc[lvl] <= subtree_max(dpad)) { L0 = lvl; break; }
    // the moving pass computes the next level's boxes when DP / 4 is a power
    // of two; boxes alternate between box and box2
    const bool fused_box = splitting_type != 0 && (dpad == 4 || dpad == 8 || dpad == 16 || dpad == 32 || dpad == 64);
    bool orig_identity = true;      // orig is still 0..n-1 (read as the index itself)
    for (int lvl = 0; lvl < L0; ++lvl) {
        LevelArgs L;
        L.nodes = 1 << lvl;
        L.cpn = host_cpn[lvl];
        L.bnd = s.bnd + host_bnd_off[lvl];
        L.node0 = L.nodes - 1;
        L.lvl = lvl;
        L.rbox = hist_rep(L.cpn);
        L.rbox_next = lvl + 1 < L0 ? hist_rep(host_cpn[lvl + 1]) : 1;
        const unsigned nblk = (unsigned)(L.nodes * L.cpn);
        uint32_t *box = (lvl & 1) ? s.box2 : s.box;
        uint32_t *nbox = (lvl & 1) ? s.box : s.box2;
        const bool need_kbox = splitting_type != 0 && (!fused_box || lvl == 0);
        if (!(fused_box && lvl + 1 < L0)) nbox = nullptr;
        const size_t hw = level_hist_words(L.nodes, L.cpn);
        const size_t hw_next = lvl + 1 < L0 ? level_hist_words(2 * L.nodes, host_cpn[lvl + 1]) : 0;
        uint32_t *H0 = s.hist, *H1 = s.hist + s.hist_words, *H2 = s.hist + 2 * s.hist_words;
        uint32_t *P1 = s.prefix, *P2 = s.prefix + s.nodes_max, *vkey = s.prefix + 2 * s.nodes_max;
        int *R1 = s.rank, *R2 = s.rank + s.nodes_max;
        if (lvl == 0 || need_kbox) {
            const size_t m = std::max((size_t)L.nodes * L.rbox * dpad, lvl == 0 ? hw : 0);
            hipLaunchKernelGGL(k_lvl_prep, dim3(blocks_for((long long)m, 256)), dim3(256), 0, st, L, dpad,
                               need_kbox ? box : (uint32_t *)nullptr, lvl == 0 ? H0 : (uint32_t *)nullptr, hw);
        }
        const int R = hist_rep(L.cpn);
        const int *oin_k = orig_identity ? (const int *)nullptr : oin;
        err = with_dp(dpad, [&](auto dp) {
            constexpr int DP = decltype(dp)::value;
            if (need_kbox)
                hipLaunchKernelGGL(k_box<DP>, dim3(nblk), dim3(kBuildTPB), 0, st, rows, oin_k, L, d, box);
            const dim3 g(nblk), b(kBuildTPB), gn((unsigned)L.nodes);
            hipLaunchKernelGGL(k_lvl_hist0<DP>, g, b, 0, st, rows, oin_k, L, d, dpad, box, splitting_type, s.nax,
                               axis, s.keys, H0, H1, H2, hw, nbox);
            hipLaunchKernelGGL(k_lvl_pick<0>, gn, b, 0, st, L, H0, R, (const uint32_t *)nullptr,
                               (const int *)nullptr, P1, R1);
            hipLaunchKernelGGL(k_lvl_pass<1>, g, b, 0, st, L, s.keys, (const uint32_t *)nullptr, H1, P1,
                               (const int *)nullptr, (uint32_t *)nullptr, (int *)nullptr);
            hipLaunchKernelGGL(k_lvl_pass<2>, g, b, 0, st, L, s.keys, H1, H2, P1, R1, P2, R2);
            hipLaunchKernelGGL(k_lvl_count, g, b, 0, st, L, s.keys, H2, P2, R2, vkey, s.cnt);
            hipLaunchKernelGGL(k_lvl_eqoff<DP>, dim3(nblk), dim3(kBuildTPB), 0, st, L, s.keys, vkey, s.cnt,
                               s.cnt_off, s.tot, s.eqoff, H0, hw_next, s.err);
            hipLaunchKernelGGL(k_move<DP>, dim3(nblk), dim3(kBuildTPB), 0, st, rows, oin_k, L, s.keys, vkey,
                               s.cnt_off, s.tot, s.eqoff, oout, s.nax, split, s.err, nbox, d);
            return hipGetLastError();
        });
        if (err != hipSuccess) return err;
        std::swap(oin, oout);
        orig_identity = false;
    }
    const float *tree_rows = nullptr;
    if (L0 < depth) {
        // the subtree blocks gather their rows through orig and write them in
        // final order (when asked: the full-patch path never reads them)
        const int *oin_k = orig_identity ? (const int *)nullptr : oin;
        err = with_dp(dpad, [&](auto dp) {
            constexpr int DP = decltype(dp)::value;
            hipLaunchKernelGGL(k_subtree<DP>, dim3(1u << L0), dim3(kSubTPB), 0, st, rows, oin_k, oin_k,
                               s.bnd + host_bnd_off[L0], L0, depth - L0, d, splitting_type,
                               need_pts ? rows_out : (float *)nullptr, oout, axis, split, s.err, leaves, pli);
            return hipGetLastError();
        });
        if (err != hipSuccess) return err;
        std::swap(oin, oout);
        orig_identity = false;
        if (need_pts) tree_rows = rows_out;
    } else {
        // no subtree pass (or no level at all): leaf bounds and point -> leaf map here
        if (orig_identity) {
            err = iota(oin, n, st);
            if (err != hipSuccess) return err;
            tree_rows = rows;   // no level moved anything: tree order is the original order
        } else if (need_pts) {
            err = with_dp(dpad, [&](auto dp) {
                constexpr int DP = decltype(dp)::value;
                hipLaunchKernelGGL(k_gather_rows<DP>, dim3(blocks_for((long long)n * (DP / 4), 256)), dim3(256), 0,
                                   st, rows, (const int *)oin, n, rows_out);
                return hipGetLastError();
            });
            if (err != hipSuccess) return err;
            tree_rows = rows_out;
        }
        orig_identity = false;
        const int nleaves = 1 << depth;
        hipLaunchKernelGGL(k_leaves, dim3(blocks_for(std::max(n, nleaves), 256)), dim3(256), 0, st,
                           s.bnd + host_bnd_off[depth], nleaves, oin, n, leaves, pli);
    }
    err = hipGetLastError();
    t.n = n;
    t.d = d;
    t.dpad = dpad;
    t.depth = depth;
    t.pts = tree_rows;
    t.orig = oin;
    t.axis = axis;
    t.split = split;
    t.leaves = leaves;
    t.pli = pli;
    return err;
}

// ---------------------------------------------------------------------------
// natural leaves and the bucketed leaf scan (phase 1)
// ---------------------------------------------------------------------------

// kdtree.c:198-214 with only_leaf_idx = 1: right iff q[axis] - split >= 0.
// The top kNatTop levels of the tree are staged in LDS and every query is
// read once into LDS, so most of the dependent steps of the descent are LDS
// accesses; the deeper levels read the tree from L2.
constexpr int kNatTop = 9;

template <int DP>
__global__ __launch_bounds__(256) void k_natural(const float *__restrict__ q, int nq, const int *__restrict__ axis,
                                                 const float *__restrict__ split, const int *__restrict__ leaves,
                                                 int depth, int m, int *leaf, int *count, int *desc) {
    __shared__ int s_axis[(1 << kNatTop) - 1];
    __shared__ float s_split[(1 << kNatTop) - 1];
    __shared__ float sq[256 * (DP + 1)];
    const int top = depth < kNatTop ? depth : kNatTop;
    const int ntop = (1 << top) - 1;
    for (int k = threadIdx.x; k < ntop; k += 256) {
        s_axis[k] = axis[k];
        s_split[k] = split[k];
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float *qs = sq + threadIdx.x * (DP + 1);
    if (i < nq) {
        float v[DP];
        load_vec<DP>(q + (size_t)i * DP, v);
#pragma unroll
        for (int j = 0; j < DP; ++j) qs[j] = v[j];
    }
    __syncthreads();
    if (i >= nq) return;
    const float *qr = q + (size_t)i * DP;
    int idx = 0;
    for (int lvl = 0; lvl < top; ++lvl) {
        const float diff = qs[s_axis[idx]] - s_split[idx];
        idx = diff >= 0.0f ? 2 * idx + 2 : 2 * idx + 1;
    }
    for (int lvl = top; lvl < depth; ++lvl) {
        const float diff = qs[axis[idx]] - split[idx];
        idx = diff >= 0.0f ? 2 * idx + 2 : 2 * idx + 1;
    }
    (void)qr;
    int l = idx - ((1 << depth) - 1);
    leaf[i] = l;
    if (desc) {
        // the leaf as its packed range (from << 1) | (size - m), as k_propleaf
        const int from = leaves[2 * l];
        desc[i] = (from << 1) | (leaves[2 * l + 1] - from - m);
    }
    if (count) wave_count1(count, l);
}

hipError_t natural_leaves(const TreeView &t, const float *q, int nq, int *leaf, int *count, int *desc,
                          hipStream_t st) {
    if (nq == 0) return hipSuccess;
    return with_dp(t.dpad, [&](auto dp) {
        constexpr int DP = decltype(dp)::value;
        hipLaunchKernelGGL(k_natural<DP>, dim3(blocks_for(nq, 256)), dim3(256), 0, st, q, nq, t.axis, t.split,
                           t.leaves, t.depth, t.n >> t.depth, leaf, count, desc);
        return hipGetLastError();
    });
}

// Blocks of the leaf scan: a leaf's bucket is cut into chunks of kScanChunk
// queries, so one crowded leaf (a flat image region) spreads over many blocks.
constexpr int kScanChunk = 256;

int leafscan_chunk() { return kScanChunk; }

// Sorted k-list of runtime length K <= KM held in registers.  The K live
// entries occupy the top slots KM-K..KM-1; the front slots hold -inf, which
// never moves, so the K-th entry is always slot KM-1 and every index is
// static.  Insertion is the reference's rule (kdtree.c:502-532): reject if
// list[K-1] <= dist, else the entry lands after every entry <= dist.
template <int KM>
__device__ __forceinline__ void list_init(float (&ld)[KM], int (&lp)[KM], int K) {
#pragma unroll
    for (int s = 0; s < KM; ++s) {
        ld[s] = (s < KM - K) ? -INFINITY : kFltMax;
        lp[s] = 0;
    }
}

template <int KM>
__device__ __forceinline__ void list_insert(float (&ld)[KM], int (&lp)[KM], float dd, int pos) {
    int r = 0;
#pragma unroll
    for (int s = 0; s < KM - 1; ++s) r += (ld[s] <= dd) ? 1 : 0;
#pragma unroll
    for (int s = KM - 1; s >= 1; --s) {
        if (s > r) { ld[s] = ld[s - 1]; lp[s] = lp[s - 1]; }
    }
#pragma unroll
    for (int s = 0; s < KM; ++s) {
        if (s == r) { ld[s] = dd; lp[s] = pos; }
    }
}

// One block per leaf: the leaf's points go to LDS, each thread scans them
// for one query of the leaf's bucket (kdtree.c:286-303, natural leaf only).
template <int DP, int KM>
__global__ __launch_bounds__(256) void k_leafscan(TreeView t, const float *__restrict__ q, const int *__restrict__ order,
                                                  const int *__restrict__ offset, const int *__restrict__ chunk_off,
                                                  int K, const int *__restrict__ propleaf,
                                                  float *__restrict__ od, int *__restrict__ op, int *__restrict__ ol) {
    extern __shared__ float4 sp4[];
    const int nleaves = 1 << t.depth;
    const int b = blockIdx.x;
    if (b >= chunk_off[nleaves]) return;
    int lo = 0, hi = nleaves - 1;     // leaf: last l with chunk_off[l] <= b
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (chunk_off[m] <= b) lo = m; else hi = m - 1;
    }
    const int leaf = lo;
    const int qb = offset[leaf] + (b - chunk_off[leaf]) * kScanChunk;
    const int qe = min(qb + kScanChunk, offset[leaf + 1]);
    if (qb >= qe) return;
    const int from = t.leaves[2 * leaf], np = t.leaves[2 * leaf + 1] - from;
    const float4 *src = reinterpret_cast<const float4 *>(t.pts + (size_t)from * DP);
    for (int c = threadIdx.x; c < np * (DP / 4); c += blockDim.x) sp4[c] = src[c];
    __syncthreads();
    const float *sp = reinterpret_cast<const float *>(sp4);
    for (int qi = qb + threadIdx.x; qi < qe; qi += blockDim.x) {
        const int qidx = order[qi];
        float qv[DP];
        load_vec<DP>(q + (size_t)qidx * DP, qv);
        float ld[KM];
        int lp[KM];
        list_init<KM>(ld, lp, K);
        for (int i = 0; i < np; ++i) {
            float dd = dist_sq<DP>(sp + i * DP, qv);
            if (ld[KM - 1] <= dd) continue;
            list_insert<KM>(ld, lp, dd, from + i);
        }
        const size_t ob = (size_t)qidx * K - (size_t)(KM - K);
#pragma unroll
        for (int s = 0; s < KM; ++s) {
            if (s >= KM - K) {
                od[ob + s] = ld[s];
                op[ob + s] = lp[s];
                if (ol) ol[ob + s] = propleaf ? propleaf[lp[s]] : 0;
            }
        }
    }
}

hipError_t leaf_scan(const TreeView &t, const float *q, int nq, const int *order, const int *offset,
                     const int *chunk_off, int K, const int *propleaf, float *out_d, int *out_p, int *out_l,
                     hipStream_t st) {
    const int nleaves = 1 << t.depth;
    const size_t lds = (size_t)t.max_leaf * t.dpad * sizeof(float);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    // chunks <= one per leaf + one per kScanChunk queries
    const unsigned nblk = (unsigned)(nleaves + (nq + kScanChunk - 1) / kScanChunk);
    return with_dp_km(t.dpad, K, [&](auto dp, auto km) {
        constexpr int DP = decltype(dp)::value;
        constexpr int KM = decltype(km)::value;
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute((const void *)k_leafscan<DP, KM>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL((k_leafscan<DP, KM>), dim3(nblk), dim3(256), lds, st, t, q, order, offset, chunk_off, K,
                           propleaf, out_d, out_p, out_l);
        return hipGetLastError();
    });
}

// ---------------------------------------------------------------------------
// wavefront k-selection
// ---------------------------------------------------------------------------

// Per-lane sorted list of (key, pos, propagated leaf), key = distance bits
// << 32 | arrival, arrivals numbered from 1.  Keys are unique, so ordering by
// key is the reference's (distance, arrival) order.  As in list_insert, the
// live entries sit in the top K slots and the front slots hold key 0, below
// every real key, so that no slot index depends on K.
template <int KM>
struct LaneList {
    uint64_t key[KM];
    int pos[KM];
    int pl[KM];

    __device__ __forceinline__ void reset(int K) {
#pragma unroll
        for (int s = 0; s < KM; ++s) {
            key[s] = (s < KM - K) ? 0ull : ~0ull;
            pos[s] = 0;
            pl[s] = 0;
        }
    }
    __device__ __forceinline__ void insert(uint64_t k, int p, int l) {
        if (k >= key[KM - 1]) return;
        int r = 0;
#pragma unroll
        for (int s = 0; s < KM - 1; ++s) r += (key[s] < k) ? 1 : 0;
#pragma unroll
        for (int s = KM - 1; s >= 1; --s) {
            if (s > r) { key[s] = key[s - 1]; pos[s] = pos[s - 1]; pl[s] = pl[s - 1]; }
        }
#pragma unroll
        for (int s = 0; s < KM; ++s) {
            if (s == r) { key[s] = k; pos[s] = p; pl[s] = l; }
        }
    }
    // smallest live key (~0 when empty)
    __device__ __forceinline__ uint64_t head() const {
        uint64_t h = ~0ull;
#pragma unroll
        for (int s = 0; s < KM; ++s) h = (key[s] != 0ull && key[s] < h) ? key[s] : h;
        return h;
    }
    // payload of the entry holding key g
    __device__ __forceinline__ void payload(uint64_t g, int &p, int &l) const {
        p = 0;
        l = 0;
#pragma unroll
        for (int s = 0; s < KM; ++s) {
            if (key[s] == g) { p = pos[s]; l = pl[s]; }
        }
    }
    // removes the smallest live entry (slot KM-K)
    __device__ __forceinline__ void pop(int K) {
#pragma unroll
        for (int s = 0; s < KM - 1; ++s) {
            if (s >= KM - K) { key[s] = key[s + 1]; pos[s] = pos[s + 1]; pl[s] = pl[s + 1]; }
        }
        key[KM - 1] = ~0ull;
    }
};

__device__ __forceinline__ uint64_t make_key(float d, uint32_t seq) {
    return ((uint64_t)__float_as_uint(d) << 32) | seq;
}

// K rounds of wave-wide minimum; lane r receives the r-th smallest entry.
template <int KM>
__device__ __forceinline__ void wave_select(LaneList<KM> &L, int K, int lane, float &od, int &opos, int &opl) {
    for (int r = 0; r < K; ++r) {
        const uint64_t mine = L.head();
        uint64_t g = mine;
#pragma unroll
        for (int m = kWave / 2; m >= 1; m >>= 1) {
            uint64_t o = shfl_xor_u64(g, m);
            g = o < g ? o : g;
        }
        const unsigned long long b = __ballot(mine == g);
        const int owner = __ffsll(b) - 1;
        int mp = 0, ml = 0;
        L.payload(g, mp, ml);
        const int wp = readlane_i(mp, owner);
        const int wl = readlane_i(ml, owner);
        if (lane == owner) L.pop(K);
        if (lane == r) {
            od = __uint_as_float((uint32_t)(g >> 32));
            opos = wp;
            opl = wl;
        }
    }
}

// ---------------------------------------------------------------------------
// exact k-NN search (row 0 and neighbors_extern), one wavefront per query
// ---------------------------------------------------------------------------

// kdtree.c:129-268: depth-first, near child first, far child iff
// (q[axis]-split)^2 <= current K-th distance; at most max_leaves leaf visits.
// phase bits per level: 0 = fresh, 1 = near child done, 2 = far child done.
template <int DP, int KM>
__global__ __launch_bounds__(256) void k_exact(TreeView t, const float *__restrict__ q, int nq, int K,
                                               long long max_leaves, const int *__restrict__ propleaf,
                                               float *__restrict__ od, int *__restrict__ op, int *__restrict__ ol) {
    const int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) / kWave));
    const int lane = threadIdx.x & (kWave - 1);
    if (w >= nq) return;
    float qv[DP];
    load_vec<DP>(q + (size_t)w * DP, qv);
    const float qlane = lane < DP ? q[(size_t)w * DP + lane] : 0.0f;
    float my_d = kFltMax;
    int my_p = 0;
    int my_l = propleaf ? propleaf[0] : 0;
    float kth = kFltMax;
    const int leaf0 = (1 << t.depth) - 1;
    int idx = 0, lvl = 0;
    uint64_t phase = 0;
    long long visited = 0;
    while (true) {
        if (lvl == t.depth) {
            if (++visited > max_leaves) break;
            const int leaf = idx - leaf0;
            const int from = t.leaves[2 * leaf], cnt = t.leaves[2 * leaf + 1] - from;
            LaneList<KM> L;
            L.reset(K);
            if (lane < K) L.insert(make_key(my_d, (uint32_t)(1 + lane)), my_p, my_l);
            for (int off = lane; off < cnt; off += kWave) {
                const int pos = from + off;
                const float dd = dist_sq<DP>(t.pts + (size_t)pos * DP, qv);
                const int pl = propleaf ? propleaf[pos] : 0;
                L.insert(make_key(dd, (uint32_t)(1 + K + off)), pos, pl);
            }
            float nd = 0.0f;
            int np = 0, nl = 0;
            wave_select<KM>(L, K, lane, nd, np, nl);
            if (lane < K) { my_d = nd; my_p = np; my_l = nl; }
            kth = readlane_f(my_d, K - 1);
            if (lvl == 0) break;
            idx = (idx - 1) >> 1;
            --lvl;
            continue;
        }
        const int ph = (int)((phase >> (2 * lvl)) & 3u);
        const float diff = readlane_f(qlane, t.axis[idx]) - t.split[idx];
        if (ph == 0) {
            phase |= (uint64_t)1 << (2 * lvl);
            idx = diff >= 0.0f ? 2 * idx + 2 : 2 * idx + 1;
            ++lvl;
        } else if (ph == 1 && diff * diff <= kth) {
            phase = (phase & ~((uint64_t)3 << (2 * lvl))) | ((uint64_t)2 << (2 * lvl));
            idx = diff >= 0.0f ? 2 * idx + 1 : 2 * idx + 2;
            ++lvl;
        } else {
            phase &= ~((uint64_t)3 << (2 * lvl));
            if (lvl == 0) break;
            idx = (idx - 1) >> 1;
            --lvl;
        }
    }
    if (lane < K) {
        od[(size_t)w * K + lane] = my_d;
        op[(size_t)w * K + lane] = my_p;
        if (ol) ol[(size_t)w * K + lane] = my_l;
    }
}

// A workgroup's candidate set in LDS: unique 64-bit (distance, arrival)
// keys with their tree positions and leaf links.  Threads append the
// candidates that beat the current K-th key (wave ballot + one LDS atomic
// per wavefront); when the set fills up it is cut back to its K smallest
// keys by rank counting (each thread ranks its items against all items with
// broadcast reads), which also orders the final list.  Keys are unique, so
// ranks are a permutation.
constexpr int kBlock = 256;
constexpr int kBatch = 4;                            // candidates per thread between checks
constexpr int kSetCap = 32 + kBatch * kBlock;        // live K <= 32 + one batch of appends
constexpr int kSetOut = 64;                          // K <= 64 kept by bset_reduce
constexpr int set_words(int cap) { return 4 * cap + 4 * kSetOut + 4; }  // key(2) pos pl | out key(2) pos pl | count
constexpr int kSetWords = set_words(kSetCap);

struct BlockSet {
    uint64_t *key;
    int *pos;
    int *pl;
    uint64_t *okey;
    int *opos;
    int *opl;
    int *count;
};

__device__ __forceinline__ BlockSet bset_at(uint32_t *base, int cap = kSetCap) {
    BlockSet S;
    S.key = reinterpret_cast<uint64_t *>(base);
    S.pos = reinterpret_cast<int *>(base + 2 * cap);
    S.pl = S.pos + cap;
    S.okey = reinterpret_cast<uint64_t *>(base + 4 * cap);
    S.opos = reinterpret_cast<int *>(base + 4 * cap + 2 * kSetOut);
    S.opl = S.opos + kSetOut;
    S.count = S.opl + kSetOut;
    return S;
}

__device__ __forceinline__ void bset_append(const BlockSet &S, bool take, uint64_t k, int p, int l) {
    const unsigned long long m = __ballot(take);
    const int cnt = __popcll(m);
    const int lane = threadIdx.x & (kWave - 1);
    int base = 0;
    if (lane == 0 && cnt) base = atomicAdd(S.count, cnt);
    base = __builtin_amdgcn_readlane(base, 0);
    if (take) {
        const int slot = base + lane_rank(m);
        S.key[slot] = k;
        S.pos[slot] = p;
        S.pl[slot] = l;
    }
}

// Number of keys among key[0, n) below kj, n >= 1; eight keys are read at a
// time (all in flight before the first compare).
__device__ __forceinline__ int rank_below(const uint64_t *key, int n, uint64_t kj) {
    int r = 0;
    for (int i = 0; i < n; i += 8) {
        uint64_t kk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) kk[u] = key[min(i + u, n - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) r += (i + u < n && kk[u] < kj) ? 1 : 0;
    }
    return r;
}

// Keeps the K smallest keys, ascending, in [0, min(n, K)).  Block-wide.
__device__ __forceinline__ void bset_reduce(const BlockSet &S, int K) {
    __syncthreads();
    const int n = *S.count;
    if (n > K) {
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const uint64_t kj = S.key[j];
            const int r = rank_below(S.key, n, kj);
            if (r < K) {
                S.okey[r] = kj;
                S.opos[r] = S.pos[j];
                S.opl[r] = S.pl[j];
            }
        }
        __syncthreads();
        if ((int)threadIdx.x < K) {
            S.key[threadIdx.x] = S.okey[threadIdx.x];
            S.pos[threadIdx.x] = S.opos[threadIdx.x];
            S.pl[threadIdx.x] = S.opl[threadIdx.x];
        }
        __syncthreads();
        if (threadIdx.x == 0) *S.count = K;
    } else if (n > 1) {
        // fewer than K+1 items: order them
        uint64_t kj = 0;
        int pj = 0, lj = 0, r = 0;
        const bool mine = (int)threadIdx.x < n;
        if (mine) {
            kj = S.key[threadIdx.x];
            pj = S.pos[threadIdx.x];
            lj = S.pl[threadIdx.x];
            r = rank_below(S.key, n, kj);
        }
        __syncthreads();
        if (mine) {
            S.key[r] = kj;
            S.pos[r] = pj;
            S.pl[r] = lj;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// Append for a one-wavefront block: the item count n is wave-uniform and
// kept in a register (no LDS atomic).
__device__ __forceinline__ void wset_append(const BlockSet &S, int &n, bool take, uint64_t k, int p, int l) {
    const unsigned long long m = __ballot(take);
    if (take) {
        const int slot = n + lane_rank(m);
        S.key[slot] = k;
        S.pos[slot] = p;
        S.pl[slot] = l;
    }
    n += __popcll(m);
}

// Minimum of v over the 64 lanes (every lane gets it): DPP within each row
// of 16 lanes, then the four rows' minima read as scalars.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0xB1, 0xF, 0xF, false));    // quad_perm 1,0,3,2
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x4E, 0xF, 0xF, false));    // quad_perm 2,3,0,1
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x141, 0xF, 0xF, false));   // row_half_mirror
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x140, 0xF, 0xF, false));   // row_mirror
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return min(min(a, b), min(c, d));
}

// bset_reduce for a one-wavefront block: with at most 64 items each lane
// holds one item, and min(n, K) rounds of a wave minimum over the 64-bit
// keys (all distinct) take them in ascending order; the items of rank < K
// move to slots 0..K-1, and lane j < K returns item j's position and leaf
// (opos/opl; other lanes keep theirs).
__device__ __forceinline__ void wset_reduce(const BlockSet &S, int &n, int K, int lane, int &opos, int &opl) {
    wave_sync();
    if (n > kWave) {
        if (lane == 0) *S.count = n;
        bset_reduce(S, K);
        n = K;
        opos = lane < K ? S.pos[lane] : 0;
        opl = lane < K ? S.pl[lane] : 0;
        return;
    }
    const bool have = lane < n;
    uint64_t mk = have ? S.key[lane] : ~0ull;
    const int mp = have ? S.pos[lane] : 0, ml = have ? S.pl[lane] : 0;
    const int m = min(n, K);
    uint64_t rk = 0;
    int rp = 0, rl = 0;
    for (int r = 0; r < m; ++r) {
        const uint32_t hi = (uint32_t)(mk >> 32), lo = (uint32_t)mk;
        const uint32_t mh = wave_min_u32(hi);
        const uint32_t mlo = wave_min_u32(hi == mh ? lo : 0xFFFFFFFFu);
        const unsigned long long win = __ballot(hi == mh && lo == mlo && mk != ~0ull);
        if (win == 0ull) break;   // no item left (not reached: r < n items hold real keys)
        const int wl = __builtin_ctzll(win);
        const int wp = __builtin_amdgcn_readlane(mp, wl), wq = __builtin_amdgcn_readlane(ml, wl);
        if (lane == r) {
            rk = ((uint64_t)mh << 32) | mlo;
            rp = wp;
            rl = wq;
        }
        if (lane == wl) mk = ~0ull;
    }
    if (lane < m) {
        S.key[lane] = rk;
        S.pos[lane] = rp;
        S.pl[lane] = rl;
        opos = rp;
        opl = rl;
    }
    wave_sync();
    n = m;
}

// Exact k-NN search as a bounded, order-free scan (one workgroup per query).
//
// The reference's depth-first search (kdtree.c:129-268) ends with the list
// of the K smallest (distance, arrival) keys among the points it visited.
// With D* the K-th smallest distance over all points, every point p with
// d(p) <= D* is visited: on p's path, a far child is skipped only if
// (q_a - s)^2 > dK(current) >= D* >= d(p) >= (q_a - p_a)^2 >= (q_a - s)^2
// (FP32 subtraction, squaring and sequential sums of non-negative terms are
// monotone), a contradiction.  So the reference's list is the K smallest keys
// among {p : d(p) <= D*}, and any set of scanned points that holds every
// point with d(p) <= D* gives the same list.  Arrival order is the
// depth-first leaf order -- near child first -- which is the integer order
// of the "went far" bits read from the root, followed by the position in the
// leaf; that pair replaces the arrival number in the key.
//
// Which leaves are scanned.  U, the K-th distance of the natural leaf, is
// >= D*.  A leaf is queued when every far turn on its path passes
// (q_a - s)^2 <= U (the reference's own test with a radius >= D*).  A
// queued leaf's cell bound is LB = sum_a g_a^2, g_a the largest |q_a - s|
// over the far turns on axis a: a point of the leaf lies on the far side of
// each of those splits (left child <= split <= right child), so
// |q_a - p_a| >= g_a and its exact distance is >= LB.  The FP32 roundings of
// LB and of d(p) (sums of at most 64 non-negative terms) are each within
// 70 units of 2^-24, so d(p) >= LB (1 - 2^-14): a leaf with
// LB (1 - 2^-14) > U' for any U' >= D* holds no list member.  Queued leaves
// are scanned in ascending LB, eight at a time; U' is the K-th distance of
// the list so far, and the scan stops at the first leaf whose bound exceeds
// it (every later one has a larger bound).
//
// Per query: near/pass bits of every internal node go to LDS (coalesced
// passes over the node arrays); every leaf is tested against its path; the
// passing leaves are queued in LDS, bounded, sorted (bitonic) and scanned by
// the whole workgroup, one leaf per wavefront per step.
constexpr int kExactQueue = 1024;
constexpr int kNodeUnroll = 8;
constexpr float kCellSafe = 1.0f - 0x1p-14f;

template <int DP>
__global__ __launch_bounds__(kBlock) void k_exact_bound(TreeView t, const float *__restrict__ q, int nq, int K,
                                                        const int *__restrict__ propleaf, float *__restrict__ od,
                                                        int *__restrict__ op, int *__restrict__ ol, int offbits,
                                                        int words) {
    extern __shared__ uint32_t smem_u[];
    const int tid = threadIdx.x, wid = tid / kWave, lane = tid & (kWave - 1);
    const BlockSet S = bset_at(smem_u);
    uint32_t *nearb = smem_u + kSetWords;
    uint32_t *passb = nearb + words;
    uint64_t *qkey = reinterpret_cast<uint64_t *>(passb + words);   // (bound bits << 32) | queue slot
    int *qleaf = reinterpret_cast<int *>(qkey + kExactQueue);
    int *qrank = qleaf + kExactQueue;
    float *qs = reinterpret_cast<float *>(qrank + kExactQueue);     // the query, for indexing by axis
    int *qn_p = reinterpret_cast<int *>(qs + DP);
    int *nv_p = qn_p + 1;
    const int w = blockIdx.x;
    if (w >= nq) return;
    float qv[DP];
    load_vec<DP>(q + (size_t)w * DP, qv);
    const float qlane = lane < DP ? q[(size_t)w * DP + lane] : 0.0f;
    if (tid < DP) qs[tid] = q[(size_t)w * DP + tid];
    const int nn = (1 << t.depth) - 1;
    const int nleaves = 1 << t.depth;

    // near bits: right child first iff q[axis] - split >= 0; pass bits: far
    // child allowed iff (q[axis] - split)^2 <= u.  kNodeUnroll groups of
    // kBlock nodes a step, their loads issued together.
    auto node_bits = [&](uint32_t *bits, bool pass, float u) {
        for (int base = 0; base < nn; base += kNodeUnroll * kBlock) {
            int a[kNodeUnroll];
            float sv[kNodeUnroll];
#pragma unroll
            for (int g = 0; g < kNodeUnroll; ++g) {
                const int n = base + g * kBlock + tid;
                a[g] = n < nn ? t.axis[n] : 0;
                sv[g] = n < nn ? t.split[n] : 0.0f;
            }
#pragma unroll
            for (int g = 0; g < kNodeUnroll; ++g) {
                const int n = base + g * kBlock + tid;
                const float diff = __shfl(qlane, a[g], kWave) - sv[g];
                const unsigned long long b = __ballot(n < nn && (pass ? diff * diff <= u : diff >= 0.0f));
                const int wb = base + g * kBlock + wid * kWave;
                if (lane == 0 && wb < nn) {
                    bits[wb >> 5] = (uint32_t)b;
                    bits[(wb >> 5) + 1] = (uint32_t)(b >> 32);
                }
            }
        }
    };
    node_bits(nearb, false, 0.0f);
    if (tid == 0) *S.count = 0;
    __syncthreads();

    // radius U: K-th distance in the natural leaf (FLT_MAX if it holds < K)
    int node = 0;
    for (int lev = 0; lev < t.depth; ++lev) node = 2 * node + 1 + (int)((nearb[node >> 5] >> (node & 31)) & 1u);
    int from, to;
    leaf_range(node - nn, t.n, t.depth, from, to);
    bset_append(S, tid < K, make_key(kFltMax, (uint32_t)tid), 0, 0);
    __syncthreads();
    for (int ob = 0; ob < to - from; ob += kBlock) {
        const int off = ob + tid;
        const bool val = off < to - from;
        const float dd = val ? dist_sq<DP>(t.pts + (size_t)(from + off) * DP, qv) : 0.0f;
        bset_append(S, val, make_key(dd, (uint32_t)(K + off)), 0, 0);
        bset_reduce(S, K);
    }
    const float U = __uint_as_float((uint32_t)(S.key[K - 1] >> 32));
    __syncthreads();

    node_bits(passb, true, U);
    // the reference's initial list: K sentinels, arrivals 0..K-1
    const int pl0 = propleaf ? propleaf[0] : 0;
    if (tid < K) {
        S.key[tid] = make_key(kFltMax, (uint32_t)tid);
        S.pos[tid] = 0;
        S.pl[tid] = pl0;
    }
    if (tid == 0) { *S.count = K; *qn_p = 0; }
    __syncthreads();
    const uint64_t ubound = (U < kFltMax) ? (((uint64_t)__float_as_uint(U) << 32) | 0xFFFFFFFFull) : ~0ull;
    uint64_t bkey = ubound < S.key[K - 1] ? ubound : S.key[K - 1];
    float ucur = __uint_as_float((uint32_t)(bkey >> 32));   // >= D*: the K-th distance so far, at most U
    const uint32_t seq0 = (uint32_t)K;

    // the queued leaves [0, count): cell bounds, sort, scan in ascending bound
    auto process_queue = [&](int count) {
        if (tid == 0) *nv_p = 0;
        __syncthreads();
        for (int j = tid; j < count; j += kBlock) {
            const int l = qleaf[j];
            int ax[16];
            float sp[16];
#pragma unroll
            for (int lev = 0; lev < 16; ++lev) {   // depth <= 15: the path's nodes, loads issued together
                const int nd = lev < t.depth ? (1 << lev) - 1 + (l >> (t.depth - lev)) : 0;
                ax[lev] = lev < t.depth ? t.axis[nd] : 0;
                sp[lev] = lev < t.depth ? t.split[nd] : 0.0f;
            }
            float gap[DP];
#pragma unroll
            for (int jj = 0; jj < DP; ++jj) gap[jj] = 0.0f;
            int rank = 0;
#pragma unroll
            for (int lev = 0; lev < 16; ++lev) {
                if (lev < t.depth) {
                    const int dir = (l >> (t.depth - 1 - lev)) & 1;
                    const float diff = qs[ax[lev]] - sp[lev];
                    const int far = dir ^ (int)(diff >= 0.0f);
                    rank = (rank << 1) | far;
                    const float g = far ? fabsf(diff) : 0.0f;
#pragma unroll
                    for (int jj = 0; jj < DP; ++jj) gap[jj] = (jj == ax[lev]) ? fmaxf(gap[jj], g) : gap[jj];
                }
            }
            float lb = 0.0f;
#pragma unroll
            for (int jj = 0; jj < DP; ++jj) lb = lb + gap[jj] * gap[jj];
            const float lbs = lb * kCellSafe;
            const bool keep = !(lbs > ucur);
            qkey[j] = keep ? (((uint64_t)__float_as_uint(lbs) << 32) | (uint32_t)j) : ~0ull;
            qrank[j] = rank;
            const unsigned long long m = __ballot(keep);
            if (lane_rank(m) == 0 && keep) atomicAdd(nv_p, __popcll(m));
        }
        int P = 1;
        while (P < count) P <<= 1;
        for (int j = count + tid; j < P; j += kBlock) qkey[j] = ~0ull;
        // bitonic sort of qkey[0, P)
        for (int k = 2; k <= P; k <<= 1)
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                __syncthreads();
                for (int i = tid; i < P; i += kBlock) {
                    const int ixj = i ^ jj;
                    if (ixj > i) {
                        const uint64_t x = qkey[i], y = qkey[ixj];
                        if ((x > y) == ((i & k) == 0)) {
                            qkey[i] = y;
                            qkey[ixj] = x;
                        }
                    }
                }
            }
        __syncthreads();
        const int nvalid = *nv_p;
        // eight leaves a step: wavefront wid takes sorted leaves j0 + wid and
        // j0 + wid + 4; each lane keeps up to kBatch points in flight
        for (int j0 = 0; j0 < nvalid; j0 += 4 * (kBatch / 2)) {
            if (__uint_as_float((uint32_t)(qkey[j0] >> 32)) > ucur) break;   // every later leaf is farther
            int lf[kBatch / 2], lc[kBatch / 2];
            uint32_t lr[kBatch / 2];
#pragma unroll
            for (int h = 0; h < kBatch / 2; ++h) {
                const int j = j0 + wid + 4 * h;
                const bool jv = j < nvalid;
                const int slot = jv ? (int)(uint32_t)qkey[j] : 0;
                int f = 0, e = 0;
                if (jv) leaf_range(qleaf[slot], t.n, t.depth, f, e);
                lf[h] = f;
                lc[h] = e - f;
                lr[h] = jv ? ((uint32_t)qrank[slot] << offbits) : 0u;
            }
            for (int ob = 0; ob < t.max_leaf; ob += 2 * kWave) {
                float pt[kBatch][DP];
                int pos[kBatch];
                bool val[kBatch];
                uint32_t sq[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int off = ob + lane + (u & 1) * kWave;
                    val[u] = off < lc[u >> 1];
                    pos[u] = lf[u >> 1] + (val[u] ? off : 0);
                    sq[u] = seq0 + (lr[u >> 1] | (uint32_t)off);
                    if (val[u]) load_vec<DP>(t.pts + (size_t)pos[u] * DP, pt[u]);
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    float s2 = 0.0f;
                    if (val[u]) {
#pragma unroll
                        for (int jj = 0; jj < DP; ++jj) {
                            const float dd = pt[u][jj] - qv[jj];
                            s2 = s2 + dd * dd;
                        }
                    }
                    const uint64_t key = make_key(s2, sq[u]);
                    const bool take = val[u] && key < bkey;
                    bset_append(S, take, key, pos[u], (take && propleaf) ? propleaf[pos[u]] : 0);
                }
                __syncthreads();
                if (*S.count > kSetCap - kBatch * kBlock) {
                    bset_reduce(S, K);
                    bkey = ubound < S.key[K - 1] ? ubound : S.key[K - 1];
                }
            }
            // the bound after this step
            bset_reduce(S, K);
            bkey = ubound < S.key[K - 1] ? ubound : S.key[K - 1];
            ucur = __uint_as_float((uint32_t)(bkey >> 32));
        }
        __syncthreads();
    };
    // every leaf against its path, two leaves per thread a step
    for (int lb = 0; lb < nleaves; lb += 2 * kBlock) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int l = lb + h * kBlock + tid;
            bool ok = l < nleaves;
            int nd = 0;
            for (int lev = t.depth - 1; lev >= 0; --lev) {
                const int dir = (l >> lev) & 1;
                const int nr = (int)((nearb[nd >> 5] >> (nd & 31)) & 1u);
                const int pb = (int)((passb[nd >> 5] >> (nd & 31)) & 1u);
                ok = ok && (dir == nr || pb);
                nd = 2 * nd + 1 + dir;
            }
            const unsigned long long m = __ballot(ok);
            const int cnt = __popcll(m);
            int base = 0;
            if (lane == 0 && cnt) base = atomicAdd(qn_p, cnt);
            base = __builtin_amdgcn_readlane(base, 0);
            if (ok) qleaf[base + lane_rank(m)] = l;
        }
        __syncthreads();
        const int qn = *qn_p;
        if (qn > kExactQueue - 2 * kBlock) {
            process_queue(qn);
            if (tid == 0) *qn_p = 0;
            __syncthreads();
        }
    }
    __syncthreads();
    process_queue(*qn_p);
    bset_reduce(S, K);
    if (tid < K) {
        od[(size_t)w * K + tid] = __uint_as_float((uint32_t)(S.key[tid] >> 32));
        op[(size_t)w * K + tid] = S.pos[tid];
        if (ol) ol[(size_t)w * K + tid] = S.pl[tid];
    }
}

hipError_t exact_search(const TreeView &t, const float *q, int nq, int K, long long max_leaves,
                        const int *propleaf, float *out_d, int *out_p, int *out_l, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    // the bounded scan needs an unlimited visit budget (a finite max_leaves
    // stops the reference's traversal part-way) and keys that fit 32 bits
    int offbits = 0;
    while ((1 << offbits) <= t.max_leaf) ++offbits;
    const bool bounded = max_leaves >= (1LL << t.depth) && t.depth <= 15 && t.depth + offbits <= 30 &&
                         K <= 32;
    if (bounded) {
        const int nn = (1 << t.depth) - 1;
        const int words = 2 * ((nn + kWave - 1) / kWave) + 2;
        const size_t lds = (size_t)(kSetWords + 2 * words + 4 * kExactQueue + kWave + 2) * sizeof(uint32_t);
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        return with_dp(t.dpad, [&](auto dp) {
            constexpr int DP = decltype(dp)::value;
            hipError_t e = hipFuncSetAttribute((const void *)k_exact_bound<DP>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_exact_bound<DP>), dim3(nq), dim3(kBlock), lds, st, t, q, nq, K, propleaf, out_d,
                               out_p, out_l, offbits, words);
            return hipGetLastError();
        });
    }
    return with_dp_km(t.dpad, K, [&](auto dp, auto km) {
        constexpr int DP = decltype(dp)::value;
        constexpr int KM = decltype(km)::value;
        hipLaunchKernelGGL((k_exact<DP, KM>), dim3(blocks_for((long long)nq * kWave, 256)), dim3(256), 0, st, t, q,
                           nq, K, max_leaves, propleaf, out_d, out_p, out_l);
        return hipGetLastError();
    });
}

// ---------------------------------------------------------------------------
// propagation (rows >= 1)
// ---------------------------------------------------------------------------

// propleaf[pos] describes the leaf of the patch one row below the B patch
// at tree position pos (base.c:265-276: o = orig[pos]; o += n_cols if
// o + n_cols < n_a; leaf = pli[o]).  Every leaf holds m = n >> depth or m + 1
// points, so the leaf is stored as its range: (from << 1) | (size - m).  The
// range identifies the leaf, and the row sweep needs no lookup to scan it.
__global__ void k_propleaf(const int *orig, const int *pli, const int *leaves, int n_b, int n_a, int n_cols, int m,
                           int *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_b) return;
    int o = orig[i];
    if (o + n_cols < n_a) o += n_cols;
    const int l = pli[o];
    const int from = leaves[2 * l];
    out[i] = (from << 1) | (leaves[2 * l + 1] - from - m);
}

hipError_t prop_leaves(const TreeView &t, int n_a, int n_cols, int *propleaf, hipStream_t st) {
    hipLaunchKernelGGL(k_propleaf, dim3(blocks_for(t.n, 256)), dim3(256), 0, st, t.orig, t.pli, t.leaves, t.n, n_a,
                       n_cols, t.n >> t.depth, propleaf);
    return hipGetLastError();
}

// One wavefront per image column, walking rows 1..n_rows-1.  The list of
// row y starts from the natural-leaf list (arrivals 1..K) and then takes the
// candidates of the K propagated leaves in order (base.c:255-304), with
// arrival 1 + K + k * stride + offset for leaf k.  Three exact shortcuts keep
// the work per row to about two leaves:
//   * only candidates strictly closer than the current K-th distance can
//     enter (an equal distance arrives later and loses), so they are filtered
//     in registers and the survivors appended to the wavefront's LDS set,
//     whose K smallest keys K rounds of a wave minimum then pick;
//   * a leaf listed several times is scored once and every surviving point
//     is appended once per occurrence, each with its own arrival (the
//     reference's duplicates);
//   * a propagated leaf equal to the query's natural leaf is not rescanned:
//     its points closer than the natural list's K-th distance are exactly
//     the natural list's entries, which are re-appended with that leaf's
//     arrivals.
// A row's unique leaves are cut into chunks of 64 points, listed in an LDS
// table (point range, owner entry, offset in the leaf); the loads of UNR
// chunks (every valid point's whole row and leaf descriptor) are issued
// together, so a row whose unique leaves fit UNR chunks (six for d <= 16; a
// 1080p row with K = 5 has five on average) costs one dependent memory round
// trip after its table.  (A
// first round of dimensions 0-3 that filtered which rows load the rest cut
// the bytes but added a round trip per row: about 5 % slower.)  Leaf ranges
// come from the leaf descriptors (propleaf), and the next
// row's query, natural leaf and natural list are requested while the current
// row is scored.  Consecutive workgroups go to different XCDs, so the
// columns are dealt out in bands: XCD i walks columns [i B, (i + 1) B), and
// neighbouring columns, whose propagated leaves overlap, share an L2.
constexpr int kSweepCap = 64 + 4 * kWave;   // K <= 64 live + four wavefronts of appends
constexpr int kSweepTab = 256;              // chunk table entries (a window of a row's chunks)
constexpr int kRowBatch = 16;             // rows whose results are stored together

template <int DP, int UNR>
__global__ __launch_bounds__(kWave) void k_sweep(TreeView t, const float *__restrict__ q, int n_rows, int n_cols,
                                                 int K, const int *__restrict__ propleaf,
                                                 const float *__restrict__ p1d, const int *__restrict__ p1p,
                                                 const int *__restrict__ p1l, const int *__restrict__ natleaf,
                                                 int *__restrict__ fl_pos, const int *__restrict__ row0_l, int stride,
                                                 int band) {
    extern __shared__ uint32_t smem_w[];
    const int lane = threadIdx.x;
    const BlockSet S = bset_at(smem_w, kSweepCap);
    int4 *tab = reinterpret_cast<int4 *>(smem_w + set_words(kSweepCap));
    int *rowbuf = reinterpret_cast<int *>(tab + kSweepTab);   // kRowBatch rows of K results
    const int x = (int)(blockIdx.x % kXcds) * band + (int)(blockIdx.x / kXcds);
    if (x >= n_cols) return;
    const int mlo = t.n >> t.depth;
    const int chmax = (mlo + kWave) / kWave;   // chunks of the largest leaf (mlo + 1 points)
    const bool act = lane < K;
    int my_l = act ? row0_l[(size_t)x * K + lane] : 0;
    float qv[DP];
    float cd = 0.0f;
    int cp = 0, cl = 0, nat = 0;
    if (n_rows > 1) {
        const size_t qi = (size_t)n_cols + x;
        load_vec<DP>(q + qi * DP, qv);
        nat = natleaf[qi];
        if (act) { cd = p1d[qi * K + lane]; cp = p1p[qi * K + lane]; cl = p1l[qi * K + lane]; }
    }
    // rows [ybase, y) wait in rowbuf; their stores go out together
    int ybase = 1;
    auto flush_rows = [&](int y_end) {
        wave_sync();
        const int cnt = (y_end - ybase) * K;
        for (int i = lane; i < cnt; i += kWave) {
            const int r = i / K, k = i - r * K;
            fl_pos[((size_t)(ybase + r) * n_cols + x) * K + k] = rowbuf[i];
        }
        ybase = y_end;
    };
    for (int y = 1; y < n_rows; ++y) {
        const size_t qi = (size_t)y * n_cols + x;
        // occ: the entries (lanes) listing this lane's leaf; the first of
        // them scans it, unless it is the natural leaf
        uint64_t occ = 0;
        for (int j = 0; j < K; ++j) {
            const int lj = readlane_i(my_l, j);
            const uint64_t b = __ballot(act && my_l == lj);
            if (lane == j) occ = b;
        }
        const bool is_nat = act && my_l == nat;
        const bool uniq = act && (occ & ((1ull << lane) - 1ull)) == 0ull && !is_nat;
        const unsigned long long natmask = __ballot(is_nat);
        const int f = uniq ? (my_l >> 1) : 0;
        const int c = uniq ? (mlo + (my_l & 1)) : 0;
        const int ch = (c + kWave - 1) / kWave;
        // exclusive prefix of ch over the lanes and its total, bit by bit
        int chpre = 0, nch = 0;
        for (int b = 0; (1 << b) <= chmax; ++b) {
            const unsigned long long m = __ballot((ch >> b) & 1);
            chpre += lane_rank(m) << b;
            nch += __popcll(m) << b;
        }
        const int nat_from = nat >> 1;
        // the natural list as the first K items
        if (act) {
            S.key[lane] = make_key(cd, (uint32_t)(1 + lane));
            S.pos[lane] = cp;
            S.pl[lane] = cl;
        }
        int n = K;
        float bound = readlane_f(cd, K - 1);
        // next row's inputs
        float nqv[DP];
        float nd0 = 0.0f;
        int np0 = 0, nl0 = 0, nnat = 0;
        const bool more = y + 1 < n_rows;
        bool fetched = false;
        bool touched = false;   // any append this row (the list may have changed)
        for (int w0 = 0; w0 < nch; w0 += kSweepTab) {
            const int wn = min(nch - w0, kSweepTab);
            if (uniq)
                for (int cc = max(chpre, w0); cc < min(chpre + ch, w0 + wn); ++cc) {
                    const int o = (cc - chpre) * kWave;
                    tab[cc - w0] = make_int4(f + o, min(kWave, c - o), lane, o);
                }
            wave_sync();
            for (int c0 = 0; c0 < wn; c0 += UNR) {
                // the chunks' table entries, then every valid point's row and
                // leaf descriptor, all issued before the first use: one
                // dependent memory round trip per window of UNR chunks
                int4 e[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) e[u] = c0 + u < wn ? tab[c0 + u] : make_int4(0, 0, 0, 0);
                int pos[UNR], meta[UNR], plc[UNR];
                bool val[UNR];
                float4 rw[UNR][DP / 4];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    val[u] = c0 + u < wn && lane < e[u].y;
                    pos[u] = val[u] ? e[u].x + lane : 0;
                    meta[u] = (e[u].z << 24) | e[u].w;   // owner entry, offset of the chunk in its leaf
                    plc[u] = 0;
                    if (c0 + u < wn) {
                        const float4 *src = reinterpret_cast<const float4 *>(t.pts + (size_t)pos[u] * DP);
#pragma unroll
                        for (int c4 = 0; c4 < DP / 4; ++c4) rw[u][c4] = src[c4];
                        plc[u] = propleaf[pos[u]];
                    }
                }
                if (!fetched && more) {
                    const size_t nqi = qi + n_cols;
                    load_vec<DP>(q + nqi * DP, nqv);
                    nnat = natleaf[nqi];
                    const int kl = act ? lane : 0;
                    nd0 = p1d[nqi * K + kl];
                    np0 = p1p[nqi * K + kl];
                    nl0 = p1l[nqi * K + kl];
                    fetched = true;
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    if (c0 + u >= wn) continue;
                    float a = 0.0f;
#pragma unroll
                    for (int c4 = 0; c4 < DP / 4; ++c4) {
                        float dd;
                        dd = rw[u][c4].x - qv[4 * c4 + 0]; a = a + dd * dd;
                        dd = rw[u][c4].y - qv[4 * c4 + 1]; a = a + dd * dd;
                        dd = rw[u][c4].z - qv[4 * c4 + 2]; a = a + dd * dd;
                        dd = rw[u][c4].w - qv[4 * c4 + 3]; a = a + dd * dd;
                    }
                    const bool take = val[u] && a < bound;
                    if (__ballot(take) == 0ull) continue;
                    touched = true;
                    const int mt = __builtin_amdgcn_readfirstlane(meta[u]);
                    const uint32_t arr0 = (uint32_t)(1 + K + (mt & 0xFFFFFF) + lane);
                    uint64_t m = readlane_u64(occ, mt >> 24);
                    while (m) {
                        const int k = __builtin_ctzll(m);
                        m &= m - 1;
                        wset_append(S, n, take, make_key(a, arr0 + (uint32_t)(k * stride)), pos[u], plc[u]);
                        if (n > kSweepCap - kWave) {
                            int dp_ = 0, dl_ = 0;
                            wset_reduce(S, n, K, lane, dp_, dl_);
                            bound = __uint_as_float((uint32_t)(S.key[K - 1] >> 32));
                        }
                    }
                }
            }
            wave_sync();   // the next window rewrites the table
        }
        if (!fetched && more) {
            const size_t nqi = qi + n_cols;
            load_vec<DP>(q + nqi * DP, nqv);
            nnat = natleaf[nqi];
            const int kl = act ? lane : 0;
            nd0 = p1d[nqi * K + kl];
            np0 = p1p[nqi * K + kl];
            nl0 = p1l[nqi * K + kl];
        }
        // propagated leaves equal to the natural leaf: re-append the natural
        // list's entries below the bound with those leaves' arrivals
        {
            uint64_t m = natmask;
            while (m) {
                const int k = __builtin_ctzll(m);
                m &= m - 1;
                const bool take = act && cd < bound;
                if (__ballot(take) == 0ull) continue;
                touched = true;
                wset_append(S, n, take, make_key(cd, (uint32_t)(1 + K + k * stride + (cp - nat_from))), cp, cl);
                if (n > kSweepCap - kWave) {
                    int dp_ = 0, dl_ = 0;
                    wset_reduce(S, n, K, lane, dp_, dl_);
                    bound = __uint_as_float((uint32_t)(S.key[K - 1] >> 32));
                }
            }
        }
        int rpos = cp, rl = cl;     // unchanged list when nothing was appended
        if (touched) wset_reduce(S, n, K, lane, rpos, rl);
        if (act) {
            rowbuf[(y - ybase) * K + lane] = rpos;
            my_l = rl;
        }
        if (y + 1 - ybase == kRowBatch || !more) flush_rows(y + 1);
        if (more) {
#pragma unroll
            for (int j = 0; j < DP; ++j) qv[j] = nqv[j];
            cd = nd0;
            cp = np0;
            cl = nl0;
            nat = nnat;
        }
        wave_sync();
    }
}

hipError_t row_sweep(const TreeView &t, const float *q, int n_rows, int n_cols, int K, const int *propleaf,
                     const float *p1d, const int *p1p, const int *p1l, const int *natleaf, int *fl_pos,
                     const int *row0_l, hipStream_t st) {
    if (n_rows < 2) return hipSuccess;
    // offsets of a leaf's chunks (below the leaf size) share a key field with the owner entry
    if (t.max_leaf >= (1 << 24)) return hipErrorInvalidValue;
    const size_t lds = (size_t)set_words(kSweepCap) * sizeof(uint32_t) + kSweepTab * sizeof(int4) +
                       (size_t)kRowBatch * kWave * sizeof(int);
    const int band = (n_cols + kXcds - 1) / kXcds;
    // chunks of the largest row: K leaves of up to max_leaf points
    const long long chunks = (long long)K * ((t.max_leaf + kWave - 1) / kWave);
    return with_dp(t.dpad, [&](auto dp) {
        constexpr int DP = decltype(dp)::value;
        auto launch = [&](auto unr) {
            constexpr int UNR = decltype(unr)::value;
            hipLaunchKernelGGL((k_sweep<DP, UNR>), dim3((unsigned)(band * kXcds)), dim3(kWave), lds, st, t, q,
                               n_rows, n_cols, K, propleaf, p1d, p1p, p1l, natleaf, fl_pos, row0_l, t.max_leaf, band);
            return hipGetLastError();
        };
        // a window holds UNR chunks' rows in registers (UNR x DP floats
        // per lane): six for DP <= 16 (a K = 5 row of 64-point leaves is one
        // window), fewer for wider points
        if (DP <= 16 && chunks > 4) return launch(std::integral_constant<int, 6>{});
        if (DP <= 16) return launch(std::integral_constant<int, 4>{});
        if (DP <= 32) return launch(std::integral_constant<int, 3>{});
        return launch(std::integral_constant<int, 2>{});
    });
}

// ---------------------------------------------------------------------------
// index conversion + best-of-K (base.c:309-353, norm_array base.c:362-374)
// ---------------------------------------------------------------------------

__device__ __forceinline__ long long patch_off(const PatchGeom &g, long long i) {
    if (g.image) {
        // window indices are below 2^31: a 32-bit division
        const unsigned ui = (unsigned)i, nc = (unsigned)g.ncols;
        const unsigned y = ui / nc, x = ui - y * nc;
        return ((long long)y * g.width + x) * 3;
    }
    return i * (long long)g.rowlen;
}

// sqrt of the sequential FP32 sum of ((float)a - (float)b)^2 (norm_array).
__device__ __forceinline__ float patch_l2_fp32(const PatchGeom &ga, long long i, const PatchGeom &gb, long long j) {
    const uint8_t *a = ga.base + patch_off(ga, i);
    const uint8_t *b = gb.base + patch_off(gb, j);
    float s = 0.0f;
    for (int r = 0; r < ga.rows; ++r) {
        const uint8_t *ra = a + r * ga.rstride;
        const uint8_t *rb = b + r * gb.rstride;
        for (int u = 0; u < ga.rowlen; ++u) {
            float df = (float)ra[u] - (float)rb[u];
            s = s + df * df;
        }
    }
    return sqrt_rn(s);
}

// 4 bytes at byte offset off of a buffer of nbytes (any alignment); bytes
// past the end read as 0 and nothing past the end is loaded.
__device__ __forceinline__ uint32_t load4(const uint8_t *base, long long off, long long nbytes) {
    const long long al = off & ~3LL;
    const int sh = (int)(off & 3);
    if (al + 8 <= nbytes) {
        const uint32_t lo = *reinterpret_cast<const uint32_t *>(base + al);
        const uint32_t hi = *reinterpret_cast<const uint32_t *>(base + al + 4);
        return __builtin_amdgcn_alignbyte(hi, lo, sh);
    }
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k)
        if (off + k < nbytes) v |= (uint32_t)base[off + k] << (8 * k);
    return v;
}

// Exact integer sums over one patch pair: sum a^2, sum b^2, sum a*b with
// 4-byte dot products (v_dot4_u32_u8); the last word of a row is masked.
__device__ __forceinline__ void patch_dots(const PatchGeom &ga, long long oa, const PatchGeom &gb, long long ob,
                                           uint32_t &saa, uint32_t &sbb, uint32_t &sab) {
    const int nw = (ga.rowlen + 3) >> 2;
    const int tail = ga.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : 0xFFFFFFFFu;
    for (int r = 0; r < ga.rows; ++r) {
        const long long ra = oa + r * ga.rstride, rb = ob + r * gb.rstride;
        for (int wd = 0; wd < nw; ++wd) {
            uint32_t wa = load4(ga.base, ra + 4 * wd, ga.nbytes);
            uint32_t wb = load4(gb.base, rb + 4 * wd, gb.nbytes);
            if (wd == nw - 1) { wa &= tmask; wb &= tmask; }
            saa = __builtin_amdgcn_udot4(wa, wa, saa, false);
            sbb = __builtin_amdgcn_udot4(wb, wb, sbb, false);
            sab = __builtin_amdgcn_udot4(wa, wb, sab, false);
        }
    }
}

// Index conversion + best-of-K (base.c:309-353).  Each candidate's squared
// distance is the exact integer S = sum a^2 + sum b^2 - 2 sum a*b.  While S
// < 2^24 every partial sum of the reference's sequential FP32 loop is an
// integer below 2^24, so that loop returns S exactly; larger sums take the
// FP32 loop itself.  The comparison is made on the correctly rounded roots,
// as in the reference (two sums can share a root).
__global__ __launch_bounds__(256) void k_best(PatchGeom ga, PatchGeom gb, const int *__restrict__ fl_pos,
                                              const int *__restrict__ orig, int nq, int K, int sel,
                                              int *__restrict__ fin, float *__restrict__ fdist, int *__restrict__ knn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int *lp = fl_pos + (size_t)i * K;
    if (knn) {
        for (int k = 0; k < K; ++k) knn[(size_t)i * K + k] = orig[lp[k]];
    }
    const long long oa = patch_off(ga, i);
    const int kk = sel ? K : 1;
    int best = -1;
    float bd = kFltMax;
    for (int k = 0; k < kk; ++k) {
        const int j = orig[lp[k]];
        if (!sel && !fdist) { best = j; break; }
        uint32_t saa = 0, sbb = 0, sab = 0;
        patch_dots(ga, oa, gb, patch_off(gb, j), saa, sbb, sab);
        const uint32_t S = saa + sbb - 2u * sab;
        const float dd = (S < (1u << 24)) ? sqrt_rn((float)S) : patch_l2_fp32(ga, i, gb, j);
        if (!sel) { best = j; bd = dd; break; }
        if (dd < bd) { bd = dd; best = j; }
    }
    fin[i] = best;
    if (fdist) fdist[i] = bd;
}

// k_best for image patches whose rows fit LW - 1 dwords and K <= kBestMax:
// the patch is walked row by row, and each row of the query and of every
// candidate is read as one LW-dword window (wide loads, all in flight
// together) instead of two dword loads per word; the integer sums and the
// selection are those of k_best.
constexpr int kBestMax = 8;

template <int LW>
__global__ __launch_bounds__(256) void k_best_w(PatchGeom ga, PatchGeom gb, const int *__restrict__ fl_pos,
                                                const int *__restrict__ orig, int nq, int K, int sel,
                                                int *__restrict__ fin, float *__restrict__ fdist,
                                                int *__restrict__ knn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int *lp = fl_pos + (size_t)i * K;
    int jb[kBestMax];
#pragma unroll
    for (int k = 0; k < kBestMax; ++k) jb[k] = k < K ? orig[lp[k]] : 0;
    if (knn)
#pragma unroll
        for (int k = 0; k < kBestMax; ++k)
            if (k < K) knn[(size_t)i * K + k] = jb[k];
    const int kk = sel ? K : 1;
    if (!sel && !fdist) {
        fin[i] = jb[0];
        return;
    }
    const int rw = (ga.rowlen + 3) >> 2;
    const int tail = ga.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : 0xFFFFFFFFu;
    const long long oa = patch_off(ga, i);
    long long ob[kBestMax];
#pragma unroll
    for (int k = 0; k < kBestMax; ++k) ob[k] = patch_off(gb, jb[k]);
    uint32_t saa = 0, sbb[kBestMax], sab[kBestMax];
#pragma unroll
    for (int k = 0; k < kBestMax; ++k) sbb[k] = sab[k] = 0;
    for (int r = 0; r < ga.rows; ++r) {
        uint32_t wa[1][LW], wb[kBestMax][1][LW];
        fetch_rows<1, LW>(ga, oa, true, r, 1, wa);
#pragma unroll
        for (int k = 0; k < kBestMax; ++k) fetch_rows<1, LW>(gb, ob[k], k < kk, r, 1, wb[k]);
        const int sha = (int)((oa + (long long)r * ga.rstride) & 3);
#pragma unroll
        for (int w = 0; w < LW - 1; ++w) {
            if (w < rw) {
                uint32_t va = __builtin_amdgcn_alignbyte(wa[0][w + 1], wa[0][w], sha);
                if (w == rw - 1) va &= tmask;
                saa = __builtin_amdgcn_udot4(va, va, saa, false);
#pragma unroll
                for (int k = 0; k < kBestMax; ++k) {
                    if (k < kk) {
                        const int shb = (int)((ob[k] + (long long)r * gb.rstride) & 3);
                        uint32_t vb = __builtin_amdgcn_alignbyte(wb[k][0][w + 1], wb[k][0][w], shb);
                        if (w == rw - 1) vb &= tmask;
                        sbb[k] = __builtin_amdgcn_udot4(vb, vb, sbb[k], false);
                        sab[k] = __builtin_amdgcn_udot4(va, vb, sab[k], false);
                    }
                }
            }
        }
    }
    int best = -1;
    float bd = kFltMax;
#pragma unroll
    for (int k = 0; k < kBestMax; ++k) {
        if (k >= kk) continue;
        const uint32_t S = saa + sbb[k] - 2u * sab[k];
        const float dd = (S < (1u << 24)) ? sqrt_rn((float)S) : patch_l2_fp32(ga, i, gb, jb[k]);
        if (!sel) { best = jb[k]; bd = dd; continue; }
        if (dd < bd) { bd = dd; best = jb[k]; }
    }
    fin[i] = best;
    if (fdist) fdist[i] = bd;
}

hipError_t best_of_k(const PatchGeom &ga, const PatchGeom &gb, const int *fl_pos, const int *orig, int nq, int K,
                     int select_best, int *final_out, float *dist_out, int *knn_out, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    const int rw = (ga.rowlen + 3) >> 2;
    auto wide = [&](auto lw) {
        constexpr int LW = decltype(lw)::value;
        hipLaunchKernelGGL((k_best_w<LW>), dim3(blocks_for(nq, 256)), dim3(256), 0, st, ga, gb, fl_pos, orig, nq, K,
                           select_best, final_out, dist_out, knn_out);
        return hipGetLastError();
    };
    if (ga.image && gb.image && K <= kBestMax && ga.rows == gb.rows && ga.rowlen == gb.rowlen) {
        if (rw <= 7) return wide(std::integral_constant<int, 8>{});
        if (rw <= 9) return wide(std::integral_constant<int, 10>{});
    }
    hipLaunchKernelGGL(k_best, dim3(blocks_for(nq, 256)), dim3(256), 0, st, ga, gb, fl_pos, orig, nq, K, select_best,
                       final_out, dist_out, knn_out);
    return hipGetLastError();
}

// neighbors_extern's conversion (base.c:117-130): original index, sqrt.
__global__ void k_convert(const int *pos, const float *d2, const int *orig, long long count, int *idx, float *dist) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    idx[i] = orig[pos[i]];
    dist[i] = sqrt_rn(d2[i]);
}

hipError_t convert_knn(const int *pos, const float *d2, const int *orig, long long count, int *idx_out,
                       float *dist_out, hipStream_t st) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_convert, dim3(blocks_for(count, 256)), dim3(256), 0, st, pos, d2, orig, count, idx_out,
                       dist_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// patch extraction fused with the PCA projection
// ---------------------------------------------------------------------------

// out[i][k] = sum_f (x_f - mean_f) * basis[k][f] for the p x p x 3 window of
// patch i (f = (r*p + c)*3 + ch, openmp/base.py:152-166), zero-padded to DP.
// Acc = float: FP32 accumulation; Acc = double: FP64, rounded once to FP32.
template <int DP, typename Acc>
__global__ __launch_bounds__(256) void k_project(const uint8_t *__restrict__ img, int h, int w, int p,
                                                 const Acc *__restrict__ basis_t, const Acc *__restrict__ mean, int d,
                                                 float *__restrict__ out) {
    extern __shared__ double smem_d[];
    Acc *sb = reinterpret_cast<Acc *>(smem_d);
    const int D = p * p * 3;
    Acc *sm = sb + (size_t)D * DP;
    for (int c = threadIdx.x; c < D * DP; c += blockDim.x) sb[c] = basis_t[c];
    for (int c = threadIdx.x; c < D; c += blockDim.x) sm[c] = mean[c];
    __syncthreads();
    const int nr = h - p + 1, nc = w - p + 1;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nr * nc) return;
    const int y = (int)(i / nc), x = (int)(i - (long long)y * nc);
    Acc acc[DP];
#pragma unroll
    for (int j = 0; j < DP; ++j) acc[j] = 0;
    const int rl = 3 * p;
    for (int r = 0; r < p; ++r) {
        const uint8_t *row = img + ((size_t)(y + r) * w + x) * 3;
        for (int u = 0; u < rl; ++u) {
            const int f = r * rl + u;
            const Acc v = (Acc)row[u] - sm[f];
            const Acc *bf = sb + (size_t)f * DP;
#pragma unroll
            for (int j = 0; j < DP; ++j) acc[j] = fma(v, bf[j], acc[j]);
        }
    }
    float4 *o4 = reinterpret_cast<float4 *>(out + (size_t)i * DP);
#pragma unroll
    for (int c = 0; c < DP / 4; ++c) {
        float4 v;
        v.x = (4 * c + 0 < d) ? (float)acc[4 * c + 0] : 0.0f;
        v.y = (4 * c + 1 < d) ? (float)acc[4 * c + 1] : 0.0f;
        v.z = (4 * c + 2 < d) ? (float)acc[4 * c + 2] : 0.0f;
        v.w = (4 * c + 3 < d) ? (float)acc[4 * c + 3] : 0.0f;
        o4[c] = v;
    }
}

// ---------------------------------------------------------------------------
// FP32 projection as an MFMA GEMM (patches x components, FP32 accumulate)
// ---------------------------------------------------------------------------
// Persistent blocks of four wavefronts hold the components (Bt[c][f], the
// patch's padded feature order, zero at the pads and for c >= d) and the mean
// in LDS; per tile of 64 patches the windows are staged as bytes (a quad of
// lanes per patch) and every wavefront multiplies its 16 patches by the
// components with v_mfma_f32_16x16x4_f32.  Lane l of a wavefront takes patch
// row l % 16 and, within each group of 16 features, the four consecutive
// features 4 (l / 16) .. +3 (one dword of bytes, one float4 of mean and of
// each component), so the sum runs over the same features as the VALU
// projection, in another order.
typedef float floatx4v __attribute__((ext_vector_type(4)));
constexpr int kProjTile = 64;

template <int DK, int CT>
__global__ __launch_bounds__(256) void k_project_mfma(PatchGeom g, long long npatch, const float *__restrict__ Bt,
                                                      const float *__restrict__ mu, int dpad, float *__restrict__ out) {
    constexpr int RB = DK + 4;   // staged row stride (bytes; an odd dword count)
    extern __shared__ __align__(16) uint8_t psm[];
    float *sBt = reinterpret_cast<float *>(psm);          // CT * 16 x DK
    float *smu = sBt + CT * 16 * DK;                       // DK
    uint8_t *X = reinterpret_cast<uint8_t *>(smu + DK);   // kProjTile x RB
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    for (int i = tid; i < CT * 16 * DK; i += 256) sBt[i] = Bt[i];
    for (int i = tid; i < DK; i += 256) smu[i] = mu[i];
    const int rw = (g.rowlen + 3) >> 2;
    const int tail = g.rowlen & 3;
    const uint32_t tmask = tail ? ((1u << (8 * tail)) - 1u) : ~0u;
    constexpr int dkw = DK / 4;
    using SS = StageShape<DK>;
    const long long ntiles = (npatch + kProjTile - 1) / kProjTile;
    const int prow = lane & 15, kq = lane >> 4;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long long p0 = tile * kProjTile;
        {
            const int pl = tid >> 2, q4 = tid & 3;
            const long long pi = p0 + pl;
            const bool pv = pi < npatch;
            const long long o = patch_origin(g, pv ? pi : 0);
            uint32_t rows[SS::RL][SS::LW];
            fetch_rows(g, o, pv, q4, 4, rows);
            uint32_t *dst = reinterpret_cast<uint32_t *>(X + (size_t)pl * RB);
            for_row_words(g, o, rows, q4, 4, rw, tmask, [&](int r, int k, uint32_t v) { dst[r * rw + k] = v; });
            for (int k = g.rows * rw + q4; k < dkw; k += 4) dst[k] = 0u;
        }
        __syncthreads();
        floatx4v acc[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[ct] = floatx4v{0.0f, 0.0f, 0.0f, 0.0f};
        const uint8_t *xr = X + (size_t)(16 * wid + prow) * RB + 4 * kq;
#pragma unroll 5
        for (int f0 = 0; f0 < DK; f0 += 16) {
            const uint32_t x4 = *reinterpret_cast<const uint32_t *>(xr + f0);
            const floatx4v m4 = *reinterpret_cast<const floatx4v *>(smu + f0 + 4 * kq);
            float a[4];
            a[0] = (float)(x4 & 255u) - m4[0];
            a[1] = (float)((x4 >> 8) & 255u) - m4[1];
            a[2] = (float)((x4 >> 16) & 255u) - m4[2];
            a[3] = (float)(x4 >> 24) - m4[3];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const floatx4v b4 = *reinterpret_cast<const floatx4v *>(sBt + (size_t)(16 * ct + prow) * DK + f0 + 4 * kq);
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b4[j], acc[ct], 0, 0, 0);
            }
        }
        // acc[ct][r]: patch 16 * wid + 4 * kq + r, component 16 * ct + prow
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int comp = 16 * ct + prow;
            if (comp < dpad) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long long pi = p0 + 16 * wid + 4 * kq + r;
                    if (pi < npatch) out[pi * dpad + comp] = acc[ct][r];
                }
            }
        }
        __syncthreads();
    }
}

// Components and mean in the MFMA kernel's padded feature order:
// Bt[c * DK + fp] and mu[fp] for padded feature fp = r * 4 rw + u, u < 3p.
__global__ void k_pack_projection(const float *__restrict__ bt32, const float *__restrict__ m32, int p, int dpad,
                                  int DK, int ctw, float *__restrict__ Bt, float *__restrict__ mu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ctw * DK) return;
    const int c = i / DK, fp = i - c * DK;
    const int rl = 3 * p, rw4 = 4 * ((rl + 3) / 4);
    const int r = fp / rw4, u = fp - r * rw4;
    const bool real = r < p && u < rl;
    const int f = r * rl + u;
    Bt[i] = (real && c < dpad) ? bt32[(size_t)f * dpad + c] : 0.0f;
    if (c == 0) mu[fp] = real ? m32[f] : 0.0f;
}

int projection_dk(int p) {
    const int rw = (3 * p + 3) / 4;
    const int need = p * rw * 4;
    return need <= 176 ? 176 : need <= 256 ? 256 : need <= 400 ? 400 : need <= 512 ? 512 : 0;
}

hipError_t project_patches(const uint8_t *img, int h, int w, int p, const float *basis_t, const float *mean_f,
                           const double *basis_t64, const double *mean64, int d, int dpad, int fp64, float *out,
                           const float *pack, int per_cu, hipStream_t st) {
    const int D = p * p * 3;
    const long long n = (long long)(h - p + 1) * (w - p + 1);
    if (n <= 0) return hipSuccess;
    return with_dp(dpad, [&](auto dp) {
        constexpr int DP = decltype(dp)::value;
        if (fp64) {
            const size_t lds = (size_t)D * (DP + 1) * sizeof(double);
            if (lds > 160 * 1024) return hipErrorInvalidValue;
            hipError_t e = hipFuncSetAttribute((const void *)k_project<DP, double>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_project<DP, double>), dim3(blocks_for(n, 256)), dim3(256), lds, st, img, h, w, p,
                               basis_t64, mean64, d, out);
        } else if (projection_dk(p) == 0 || DP > 32) {
            // p > 12 or d > 32: the VALU kernel (k_project)
            const size_t lds = (size_t)D * (DP + 1) * sizeof(float);
            if (lds > 160 * 1024) return hipErrorInvalidValue;
            hipError_t e = hipFuncSetAttribute((const void *)k_project<DP, float>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_project<DP, float>), dim3(blocks_for(n, 256)), dim3(256), lds, st, img, h, w, p,
                               basis_t, mean_f, d, out);
        } else {
            // FP32 on MFMA over the table of pack_projection
            const int DKp = projection_dk(p);
            if (!pack) return hipErrorInvalidValue;
            const int ct = DP <= 16 ? 1 : 2;
            const float *Bt = pack, *mu = pack + (size_t)16 * ct * DKp;
            const PatchGeom g{img, 1, w - p + 1, w, p, 3 * p, 3LL * w, 3LL * w * h};
            int dev = 0, ncu = 256;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            const long long ntiles = (n + kProjTile - 1) / kProjTile;
            const long long cap = (long long)(per_cu > 0 ? per_cu : 2) * ncu;
            const unsigned grid = (unsigned)(ntiles < cap ? ntiles : cap);
            auto launch = [&](auto dk, auto ctv) -> hipError_t {
                constexpr int DK = decltype(dk)::value, CT = decltype(ctv)::value;
                const size_t lds = ((size_t)CT * 16 * DK + DK) * sizeof(float) + (size_t)kProjTile * (DK + 4);
                hipError_t e = hipFuncSetAttribute((const void *)k_project_mfma<DK, CT>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e != hipSuccess) return e;
                hipLaunchKernelGGL((k_project_mfma<DK, CT>), dim3(grid), dim3(256), lds, st, g, n, Bt, mu, DP, out);
                return hipGetLastError();
            };
            auto with_ct = [&](auto dk) -> hipError_t {
                return ct == 1 ? launch(dk, std::integral_constant<int, 1>{})
                               : launch(dk, std::integral_constant<int, 2>{});
            };
            switch (DKp) {
            case 176: return with_ct(std::integral_constant<int, 176>{});
            case 256: return with_ct(std::integral_constant<int, 256>{});
            case 400: return with_ct(std::integral_constant<int, 400>{});
            default: return with_ct(std::integral_constant<int, 512>{});
            }
        }
        return hipGetLastError();
    });
}

size_t projection_pack_floats(int p, int dpad) {
    const int DK = projection_dk(p);
    return (DK == 0 || dpad > 32) ? 1 : (size_t)(dpad <= 16 ? 16 : 32) * DK + DK;
}

hipError_t pack_projection(const float *basis_t, const float *mean_f, int p, int dpad, float *pack, hipStream_t st) {
    const int DK = projection_dk(p);
    if (DK == 0 || dpad > 32) return hipSuccess;   // no table: the VALU kernel projects
    const int ctw = dpad <= 16 ? 16 : 32;
    hipLaunchKernelGGL(k_pack_projection, dim3(blocks_for((long long)ctw * DK, 256)), dim3(256), 0, st, basis_t,
                       mean_f, p, dpad, DK, ctw, pack, pack + (size_t)ctw * DK);
    return hipGetLastError();
}

}  // namespace annf
