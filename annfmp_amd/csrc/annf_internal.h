// Internal declarations shared by the kernel and ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace annf {

constexpr int kWave = 64;
constexpr float kFltMax = 3.402823466e+38f;  // list sentinel, float.h:24

// XCDs of an MI355X (workgroups are handed to them round-robin by index)
constexpr int kXcds = 8;
// points per block of the tree-level kernels (one chunk of a node)
constexpr int kBuildChunk = 2048;

// rank of this lane among the lanes set in m
__device__ __forceinline__ int lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ctr[key] += 1 per active lane: lanes sharing the wave's first two keys add
// once per key, the rest on their own; no return value, so nothing waits.
__device__ __forceinline__ void wave_count1(int *ctr, int key) {
    bool done = false;
    for (int round = 0; round < 2; ++round) {
        if (!done) {
            const int lead = __builtin_amdgcn_readfirstlane(key);
            const bool mine = key == lead;
            const unsigned long long m = __ballot(mine);
            if (mine) {
                if (lane_rank(m) == 0) atomicAdd(&ctr[lead], __popcll(m));
                done = true;
            }
        }
    }
    if (!done) atomicAdd(&ctr[key], 1);
}

// A k-d tree resident in HBM.  Tree positions are the reference's XI order
// (kdtree.c:434-451 then the in-place partitions): pts holds the reduced
// vectors in that order, zero-padded from d to dpad (a multiple of 4) so a
// point is dpad/4 float4 loads.
struct TreeView {
    int n = 0, d = 0, dpad = 0, depth = 0;
    int max_leaf = 0;                 // largest leaf population
    const float *pts = nullptr;       // n * dpad
    const int *orig = nullptr;        // n
    const int *axis = nullptr;        // 2^depth - 1
    const float *split = nullptr;     // 2^depth - 1
    const int *leaves = nullptr;      // 2 * 2^depth, [from, to)
    const int *pli = nullptr;         // n, by original index
};

// Device scratch owned by the caller of the build.
struct BuildScratch {
    float *pts_a;                     // n * dpad: the input rows, read in original order
    float *pts_b;                     // n * dpad: the rows in tree order (written at the end)
    int *orig_a, *orig_b;             // n each
    uint32_t *keys;                   // n
    int *eqoff;                       // n
    uint32_t *box, *box2;             // the largest level_box_words each (this / next level)
    int *nax;                         // nodes_max
    uint32_t *prefix;                 // 3 * nodes_max (two radix prefixes, the split key)
    int *rank;                        // 2 * nodes_max
    uint32_t *hist;                   // 3 * hist_words (the three radix passes)
    size_t hist_words;                // max over the levels of level_hist_words(nodes, cpn)
    size_t nodes_max;
    int2 *cnt, *cnt_off;              // total chunks of the widest level each
    int2 *tot;                        // nodes_max
    int *bnd;                         // all levels' bounds
    int *err;                         // 1 int, non-zero if a kernel saw an impossible state
};

// How a patch is laid out for best-of-K: `rows` runs of `rowlen` bytes,
// consecutive runs `rstride` apart.  Patch i of an image starts at
// ((i / ncols) * width + i % ncols) * 3; a patch matrix row at i * rowlen.
struct PatchGeom {
    const uint8_t *base;
    int image;        // 1: windows of an image, 0: patch matrix
    int ncols;        // patches per image row (image mode)
    int width;        // image width in pixels (image mode)
    int rows;
    int rowlen;
    long long rstride;
    long long nbytes;  // size of the buffer at base (reads never pass it)
};

// ---- launchers (annf_kernels.hip); all enqueue on `st` ----
hipError_t pad_rows(const float *src, int n, int d, int dpad, float *dst, hipStream_t st);
hipError_t iota(int *dst, int n, hipStream_t st);

// Builds the tree; `t` receives the device pointers of the result (inside
// scratch ping-pong buffers and the given node/leaf arrays).
// host_maxc[l]: the largest node population of level l.
// words of one radix-pass histogram of a level (copies included)
size_t level_hist_words(int nodes, int cpn);
// words of a level's node boxes (copies included)
size_t level_box_words(int nodes, int cpn, int dpad);
hipError_t build_tree(int n, int d, int dpad, int depth, int splitting_type,
                      const int *host_bnd, const int *host_bnd_off, const int *host_cpn,
                      const int *host_maxc, BuildScratch &s, int *axis, float *split, int *leaves, int *pli,
                      TreeView &t, hipStream_t st, bool need_pts);

// Natural leaf of every query (leaf), optional per-leaf counts, and optional
// packed leaf ranges (desc: (from << 1) | (size - (n >> depth))).
hipError_t natural_leaves(const TreeView &t, const float *q, int nq, int *leaf, int *count, int *desc,
                          hipStream_t st);
int leafscan_chunk();
// Buckets n entries by leaf (leaf < 0: left out, after every bucket) with a
// stable radix sort: order = entry indices by (leaf, index), offset[0..nleaves],
// chunk_off[0..nleaves] = exclusive sum of ceil(size / chunk) over the leaves.
// keys: n ints of scratch; scratch of sort_scratch_bytes(n, nleaves).
size_t sort_scratch_bytes(int n, int nleaves);
hipError_t bucket_sorted(const int *leaf, int n, int nleaves, int chunk, int *keys, int *order, int *offset,
                         int *chunk_off, void *scratch, size_t scratch_bytes, hipStream_t st);
hipError_t leaf_scan(const TreeView &t, const float *q, int nq, const int *order, const int *offset,
                     const int *chunk_off, int K, const int *propleaf, float *out_d, int *out_p,
                     int *out_l, hipStream_t st);
hipError_t exact_search(const TreeView &t, const float *q, int nq, int K, long long max_leaves,
                        const int *propleaf, float *out_d, int *out_p, int *out_l,
                        hipStream_t st);
hipError_t prop_leaves(const TreeView &t, int n_a, int n_cols, int *propleaf, hipStream_t st);
hipError_t row_sweep(const TreeView &t, const float *q, int n_rows, int n_cols, int K,
                     const int *propleaf, const float *p1d, const int *p1p, const int *p1l,
                     const int *natleaf, int *fl_pos, const int *row0_l, hipStream_t st);
hipError_t best_of_k(const PatchGeom &ga, const PatchGeom &gb, const int *fl_pos,
                     const int *orig, int nq, int K, int select_best, int *final_out,
                     float *dist_out, int *knn_out, hipStream_t st);
hipError_t convert_knn(const int *pos, const float *d2, const int *orig, long long count,
                       int *idx_out, float *dist_out, hipStream_t st);
// FP64: CUDA-core FMA over basis_t64 / mean64.  FP32: MFMA GEMM over `pack`,
// the padded component table pack_projection writes (projection_pack_floats
// floats); per_cu persistent blocks per CU.
hipError_t project_patches(const uint8_t *img, int h, int w, int p, const float *basis_t,
                           const float *mean_f, const double *basis_t64, const double *mean64,
                           int d, int dpad, int fp64, float *out, const float *pack, int per_cu, hipStream_t st);
size_t projection_pack_floats(int p, int dpad);
hipError_t pack_projection(const float *basis_t, const float *mean_f, int p, int dpad, float *pack, hipStream_t st);

// every leaf's points in ascending B index (the position order of the
// full-patch leaf passes)
hipError_t leaf_point_order(const TreeView &t, int *out, hipStream_t st);
// exact BF16 full-patch leaf k-NN (annf_leafbf16.hip): queries bucketed with
// chunks of leafbf16_chunk(); images as two centred BF16 copies
// (center_image, buffers of 2 * cimg_elems elements) with per-patch norms of
// x - 128 (patch_norms_centered, tmp: h * (w - p + 1) words); nprobe leaves
// leaf ^ 0 .. leaf ^ (nprobe - 1); pts_order from leaf_point_order; scratch
// of leafbf16_scratch_bytes (each block's work item).  order holds output
// slots; slot s is A patch a0 + s / slot_div and receives its K exact keys
// ((d2 << 32) | B index, ascending) in keys_out[s * K ...], or, with nn
// given, the field itself (knn, nn, dist as finish_field writes them).
int leafbf16_chunk();
hipError_t leaf_blocks_q(const int *chunk_off, const int *offset, const int *leaves, int nleaves, int nblk,
                         int4 *blk_info, hipStream_t st);
bool leafbf16_supported(int p);
size_t leafbf16_scratch_bytes(int nq, int nleaves);
long long cimg_elems(int h, int w);
hipError_t center_image(const uint8_t *img, int h, int w, uint16_t *out, hipStream_t st);
hipError_t patch_norms_centered(const uint8_t *img, int h, int w, int p, uint32_t *tmp, uint32_t *norm,
                                hipStream_t st);
hipError_t leaf_bf16(const TreeView &t, const uint16_t *ca, int wa, const uint16_t *cb, int wb, int p,
                     const uint32_t *norm_a, const uint32_t *norm_b, const int *pts_order, const int *order,
                     const int *offset, const int *chunk_off, void *scratch, int nq, int nprobe, int K, int slot_div,
                     uint64_t *keys_out, int *knn, int *nn, float *dist, hipStream_t st, int ha, int hb,
                     int a0 = 0);
// FP8 full-patch leaf k-NN (annf_leaffp8.hip): patch_fp8 writes an image's
// per-patch E4M3 codes (leaffp8_dq(p) bytes per patch), norms and inverse
// scales; leaf_fp8 scans the probed leaves of the bucketed slots (chunks of
// leafbf16_chunk(), scratch of leafbf16_scratch_bytes; pb bits hold a
// position in the largest leaf) and writes, for slot s and probe r, the S
// smallest approximate keys (d~ bits rounded down << 32 | B index,
// ascending) at keys[(s * nprobe + r) * leaffp8_stride(S)]; fp8_select
// (queries in the order qorder, a permutation of q0 .. q0 + nq - 1, or
// ascending) merges each query's nsrc lists (slots (q - q0) * nsrc + r;
// with pairleaf, slots of leaf -1 are skipped) into its S smallest, scores
// them exactly and keeps the K smallest with the query's base keys (may be
// null); norm_a / norm_b: centred patch norms (patch_norms_centered).
int leaffp8_dq(int p);
bool leaffp8_supported(int p);
int leaffp8_stride(int S);
hipError_t patch_fp8(const uint8_t *img, int h, int w, int p, uint8_t *q, float *nrm, float *inv, hipStream_t st);
hipError_t leaf_fp8(const TreeView &t, const uint8_t *qa, const float *na, const float *ia, const uint8_t *qb,
                    const float *nb, const float *ib, int p, const int *pts_order, const int *order,
                    const int *offset, const int *chunk_off, void *scratch, int nq, int nprobe, int S, int slot_div,
                    int a0, int pb, uint64_t *keys, hipStream_t st);
hipError_t fp8_select(const uint64_t *keys, int S, int q0, int nq, const int *qorder, int nsrc, const int *pairleaf,
                      const PatchGeom &ga, const PatchGeom &gb, const uint32_t *norm_a, const uint32_t *norm_b,
                      const uint64_t *base, int K, uint64_t *keys_out, int *knn, int *nn, float *dist,
                      hipStream_t st);
// propagated (query, leaf) pairs of the queries q0 .. q0 + nq - 1 from the
// keys of the row above (keys1, by query): pairleaf[(q - q0) * prop_k + k]
// (-1: none, or a leaf the first pass scanned: natleaf ^ r, r < nprobe),
// per-leaf counts
hipError_t prop_pairs(const uint64_t *keys1, int q0, int nq, int K, int prop_k, int nprobe, const int *natleaf,
                      const int *pli, int n_cols, int n_b, int *pairleaf, int *count, hipStream_t st);
// the K smallest of each query's first-pass keys (keys1, by query, each
// list ascending) and its scanned pairs' keys (keys2, by query - q0; pairs
// with pairleaf -1 are skipped, so keys2 need not be filled for them) for
// queries q0 .. q0 + nq - 1; any output may be null; keys_out: the final
// keys, by query
hipError_t finish_field(const uint64_t *keys1, const uint64_t *keys2, const int *pairleaf, int q0, int nq, int K,
                        int prop_k, int *knn, int *nn, float *dist, uint64_t *keys_out, hipStream_t st);

// learned encoder (annf_mlp.hip): parameters W1 (h1 x D), b1, W2 (h2 x h1),
// b2, W3 (d x h2), b3 in one FP32 buffer; inputs (x - 128) / 128
long long mlp_param_count(int D, int h1, int h2, int d);
size_t mlp_workspace_floats(long long n, int h1, int h2, int d);
hipError_t mlp_forward(const uint8_t *x, long long n, int D, const float *params, int h1, int h2, int d,
                       float *ws, float *out, hipStream_t st);
hipError_t mlp_triplet_grad(const uint8_t *x, int B, int D, const float *params, int h1, int h2, int d,
                            float margin, float *grads, float *loss, float *ws, hipStream_t st);
hipError_t mlp_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                    float eps, int step, hipStream_t st);
int encoder_dk(int p);
bool encoder_supported(int p, int h1, int h2, int d);
hipError_t pack_encoder(const float *params, int p, int h1, int h2, int d, __bf16 *W1p, float *b1, __bf16 *W2,
                        float *b2, __bf16 *W3p, float *b3p, hipStream_t st);
// per_cu: persistent blocks per CU (2 fill the GPU; 1 leaves room for
// kernels of another stream)
hipError_t encode_patches(const PatchGeom &g, long long npatch, const __bf16 *W1p, const float *b1,
                          const __bf16 *W2, const float *b2, const __bf16 *W3p, const float *b3p, int d, int dpad,
                          float *out, hipStream_t st, int per_cu = 2);

// exact field (annf_eval.hip)
// centred patch rows (x ^ 0x80, zero padded to dq = exact_dq(p) bytes; 0:
// unsupported) of the patches idx[0 .. n) (or 0 .. n), their norms, and the
// exact k nearest B patches of the nq query rows (part: nq * exact_splits(nb)
// * k keys of scratch)
int exact_dq(int p);
hipError_t patch_matrix(const uint8_t *img, int h, int w, int p, const int *idx, long long n, uint8_t *out, int dq,
                        hipStream_t st);
hipError_t patch_norms(const uint8_t *m, long long n, int dq, uint32_t *nrm, hipStream_t st);
int exact_splits(long long nb);
hipError_t exact_sample(const uint8_t *am, const uint32_t *an, int nq, const uint8_t *bm, const uint32_t *bn,
                        long long nb, int p, int dq, int k, uint64_t *part, int *idx, int *d2, hipStream_t st);

// FP64 PCA covariance of sampled patches (annf_eval.hip): scratch of
// pca_cov_scratch(D, na + nb) bytes (256-aligned)
size_t pca_cov_scratch(int D, int ns);
hipError_t pca_cov(const PatchGeom &ga, const int *ia, int na, const PatchGeom &gb, const int *ib, int nb, int D,
                   uint8_t *scratch, double *mean, double *C, hipStream_t st);
// d leading components (d x D row-major) and explained variances from the
// eigenvectors V (columns, D x D column-major) and ascending eigenvalues w
hipError_t pca_components(const double *V, const double *w, int D, int d, double *comp, double *var, hipStream_t st);

// The openclopt backend's semantics (annf_ocl.hip): the tree of
// buildKDtree.fut on m points, padded to mp = (q + 1) * ppl with +inf rows.
struct OclTree {
    int height = -1, q = 0, ppl = 0, mp = 0;
    uint32_t *lub = nullptr;                  // 2d: global bounds (ordered keys)
    int *indir = nullptr, *indir2 = nullptr;  // mp: tree position -> point (and sort ping-pong)
    unsigned long long *keys = nullptr, *keys2 = nullptr;   // mp sort keys
    int *dims = nullptr, *clanc = nullptr;    // q
    float *vals = nullptr;                    // q
    float *leaf_pts = nullptr;                // mp x d, rows in tree order
    int *orig2leaf = nullptr;                 // mp
};
void ocl_tree_shape(int m, int defppl, int &height, int &q, int &ppl, int &mp);
size_t ocl_sort_bytes(int mp);
hipError_t ocl_build(const float *X, int m, int d, int defppl, const OclTree &t, void *sort_tmp, size_t sort_bytes,
                     hipStream_t st);
// natural-leaf lists of every query, exact lists for row 0, the row chain:
// pos / dst are n_rows * n_cols x 8 (tree positions, squared distances)
hipError_t ocl_queries(const OclTree &t, const float *qa, int n_rows, int n_cols, int d, int *nat, int *pos,
                       float *dst, hipStream_t st);
// positions -> original indices (knn, may be null) and the best of the 8 by
// exact full-patch distance (nn, may be null; needs the images)
hipError_t ocl_best(const int *pos, const int *indir, int n_a, const uint8_t *img_a, int wa, int nca,
                    const uint8_t *img_b, int wb, int ncb, int p, int *knn, int *nn, hipStream_t st);

int dpad_of(int d);
bool dpad_supported(int dpad);
bool k_supported(int K);

}  // namespace annf
