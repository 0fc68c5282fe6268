/*
 * annf_hip.h -- C ABI of the MI355X (gfx950) propagation-assisted k-d-tree
 * nearest-neighbour-field library (annfmp_amd/libannf_hip.so).
 *
 * Each entry point replaces one function of the reference's openmp C core
 * (annfmp/openmp/kdtree/src/include/base.h, bound to Python through SWIG in
 * annfmp/openmp/kdtree/swig/cpu_float.i).  The argument order and meaning
 * follow the reference; the conventions differ where the reference's cannot
 * serve a GPU library:
 *   - state lives in an opaque handle (annf_ctx) instead of the caller-owned
 *     KD_TREE_PARAMETERS / KD_TREE_RECORD structs;
 *   - every call returns an int status (ANNF_OK or a negative ANNF_E* code,
 *     text in annf_last_error()) instead of printf + exit();
 *   - "host" entry points take host buffers and copy them; "_dev" entry
 *     points take device pointers (HBM-resident data, no copies);
 *   - every call runs on the handle's own HIP stream and returns after that
 *     stream has finished, unless its name ends in _async;
 *   - floating-point types are fixed to float (the reference's USE_DOUBLE=0
 *     build, the only one its Python front ever selects).
 *
 * Threading: one handle must not be used from two threads at once; distinct
 * handles are independent.
 */
#ifndef ANNF_HIP_H
#define ANNF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANNF_OK 0
#define ANNF_EINVAL (-1)       /* bad argument (the reference exits here)  */
#define ANNF_ENOTREE (-2)      /* no tree fitted on this handle            */
#define ANNF_EHIP (-3)         /* HIP runtime error                        */
#define ANNF_ENOMEM (-4)       /* device allocation failed                 */
#define ANNF_EUNSUPPORTED (-5) /* outside this build's limits (K, d)       */

#define ANNF_SPLIT_CYCLIC 0    /* global.h:14 SPLITTING_TYPE_CYCLIC         */
#define ANNF_SPLIT_LONGEST 1   /* global.h:15 SPLITTING_TYPE_LONGEST_BOX    */

/* Build limits: n_neighbors <= ANNF_MAX_K (the full-patch entries
 * annf_field_mm_dev: <= 32), reduced dimension <= ANNF_MAX_D (beyond them:
 * ANNF_EUNSUPPORTED; the reference accepts n_neighbors up to 1000,
 * helper.c:32, and any dimension). */
#define ANNF_MAX_K 64
#define ANNF_MAX_D 64

typedef struct annf_ctx annf_ctx;

/* ---- handle management (no reference counterpart: the reference keeps
 *      global state and never frees it, base.c:379-381) ---------------- */
int annf_create(annf_ctx **out, int device);
int annf_destroy(annf_ctx *ctx);
const char *annf_last_error(void);
/* The handle's HIP stream, as a hipStream_t cast to void*. */
void *annf_stream(annf_ctx *ctx);

/* ---- init_extern (base.h:36-42) --------------------------------------
 * Validates and stores the parameters (check_parameters, helper.c:30-52).
 * num_threads is accepted for signature parity and ignored. */
int annf_init(annf_ctx *ctx, int n_neighbors, int tree_depth, int max_leaves,
              int num_threads, int splitting_type, int verbosity_level);

/* ---- fit_extern (base.h:53-57) ---------------------------------------
 * Builds the tree over Xtrain (nXtrain x dXtrain, row-major float32) at the
 * depth given to annf_init.  The tree stays on the device. */
int annf_fit(annf_ctx *ctx, const float *Xtrain, int nXtrain, int dXtrain);
int annf_fit_dev(annf_ctx *ctx, const float *Xtrain_dev, int nXtrain, int dXtrain);

/* ---- neighbors_extern (base.h:76-86) ---------------------------------
 * Exact k nearest neighbours (backtracking search, at most max_leaves leaf
 * visits) of nXtest queries; K = ddistances = dindices.  Writes sqrt
 * distances and original training indices. */
int annf_neighbors(annf_ctx *ctx, const float *Xtest, int nXtest, int dXtest,
                   float *distances, int ndistances, int ddistances,
                   int *indices, int nindices, int dindices);

/* ---- process_rows_extern (base.h:89-107) -----------------------------
 * Row 0: exact search; rows >= 1: natural leaf + (propagate) the leaves of
 * the patches one row below the K matches of the patch above; then index
 * conversion and (select_best_nn) best-of-K by full-patch L2 on patchesA /
 * patchesB.  patchesA/patchesB are only read when select_best_nn != 0. */
int annf_process_rows(annf_ctx *ctx, int *finalindices, int nfinalindices,
                      const uint8_t *patchesA, int npatchesA, int dpatchesA,
                      const uint8_t *patchesB, int npatchesB, int dpatchesB,
                      const float *patchesred, int npatchesred, int dpatchesred,
                      int n_rows, int n_cols, int n_neighbors, int propagate,
                      int select_best_nn);

/* ---- norm_array (base.h:109) -----------------------------------------
 * Host helper with the reference's arithmetic: sqrt of the sequential FP32
 * sum of squared differences of two uint8 vectors. */
float annf_norm_array(const uint8_t *patchA, const uint8_t *patchB, int dim);

/* ---- free_resources_extern (base.h:115) ------------------------------
 * Releases the tree and all device workspaces of the handle. */
int annf_free_resources(annf_ctx *ctx);

/* ---- tree export (no reference counterpart; the reference's arrays are
 *      reachable through its KD_TREE_RECORD, global.h:38-50) -------------
 * Any output pointer may be NULL.  Sizes: axis/split 2^depth-1, leaves
 * 2*2^depth, xi_orig n, xi_pts n*d, pli n. */
int annf_tree_info(annf_ctx *ctx, int *n, int *d, int *depth);
int annf_tree_export(annf_ctx *ctx, int *axis, float *split, int *leaves,
                     int *xi_orig, float *xi_pts, int *pli);
/* The inverse of annf_tree_export (no reference counterpart: the reference
 * refits on every process start, kdtree/base.py:92-153).  Restores a tree
 * for annf_neighbors on a handle set up by annf_init with the same depth.
 * Host arrays, every one required, sizes as above.  Checked before anything
 * is uploaded: the leaf ranges must be the median split of n points over
 * depth levels, xi_orig a permutation of [0, n), pli the leaf of each
 * original index, every axis in [0, d) (ANNF_EINVAL otherwise; a failed
 * import leaves the handle without a tree). */
int annf_tree_import(annf_ctx *ctx, int n, int d, int depth, const int *axis,
                     const float *split, const int *leaves, const int *xi_orig,
                     const float *xi_pts, const int *pli);

#define ANNF_PROJ_PCA32 0
#define ANNF_PROJ_PCA64 1
#define ANNF_PROJ_ENCODER 2

/* ---- image-level fused field (the openmp backend's fit,
 *      annfmp/openmp/base.py:53-150, from projection onwards) -----------
 * All pointers are device pointers.  img_a is ha x wa x 3, img_b is
 * hb x wb x 3 (uint8, row-major).  basis is d x (p*p*3) (PCA components,
 * sklearn's components_ layout): float32 for ANNF_PROJ_PCA32, float64 for
 * ANNF_PROJ_PCA64; mean is p*p*3 float64.
 * Patches are projected as sum_f (x_f - mean_f) * basis[k][f], accumulated
 * in FP32 (projection = ANNF_PROJ_PCA32) or FP64 (ANNF_PROJ_PCA64, then
 * rounded to FP32: the reference's FP64 transform, openmp/base.py:193-200,
 * up to its summation order), or by the handle's learned encoder
 * (ANNF_PROJ_ENCODER,
 * set with annf_set_encoder; basis and mean may then be NULL).
 * The tree over B's projected patches is built at depth tree_depth.
 * Outputs: nn_out (n_a int32, original B indices), optionally dist_out
 * (n_a float32: full-patch L2 of the chosen match), knn_out (n_a x K int32
 * original indices before best-of-K), red_a_out / red_b_out (n x d
 * float32 projected patches).  Any optional output may be NULL. */
int annf_field_dev(annf_ctx *ctx, const uint8_t *img_a, int ha, int wa,
                   const uint8_t *img_b, int hb, int wb, int psize,
                   const void *basis, const double *mean, int d,
                   int projection, int tree_depth, int n_neighbors,
                   int propagate, int select_best_nn, int *nn_out,
                   float *dist_out, int *knn_out, float *red_a_out,
                   float *red_b_out);

/* ---- precision variants of the field (no reference counterpart: the
 *      reference's leaf scan, kdtree.c:305-352, measures distances in the
 *      projected space) ----------------------------------------------------
 * Same projection (an ANNF_PROJ_* value) and tree as annf_field_dev.
 * Every A patch is routed to its natural leaf; the points of the leaves
 * leaf ^ 0 .. leaf ^ (nprobe - 1) (nprobe a power of two: the enclosing
 * subtree of nprobe leaves) are compared over the full p*p*3 patch as
 * |a|^2 + |b|^2 - 2 a.b on MFMA with FP32 accumulation:
 *   ANNF_PREC_BF16: BF16 operands; the distances are exact integers, the
 *     K list is the K smallest (distance, B index);
 *   ANNF_PREC_FP8:  FP8-E4M3 operands (annf_patch_fp8_dev's codes, on
 *     v_mfma_f32_32x32x64_f8f6f4); each scanned (patch, leaf) keeps its
 *     `shortlist` smallest approximate distances (n_neighbors <= shortlist
 *     <= 32, nprobe <= 32), each A patch's `shortlist` smallest over its
 *     scanned leaves are re-ranked with exact integer distances into the K
 *     list.
 * propagate = m > 0 adds the reference's propagated leaves (base.c:255-282)
 * as a second, parallel pass: for a patch below the first row, the leaves of
 * the B patches one row below the first min(m, K) first-pass matches of the
 * patch above (m >= K: all of them, the reference's rule; dropping its
 * natural leaf and repeats) are scanned the same way, and the K smallest
 * keys of both passes are kept.  The candidates come from the
 * first pass of the row above, not from its propagated result (one parallel
 * pass); propagate = -m < 0 (ANNF_PREC_BF16 only) is the reference's row
 * chain instead: rows in order, each taking its candidates from the final
 * lists of the row above.  With ANNF_PREC_FP8 the propagated leaves' lists
 * form one shortlist per A patch, re-ranked exactly and merged with its
 * first-pass list.  Either needs wa == wb.  A propagated leaf
 * the first pass already scanned (natural leaf ^ r, r < nprobe) is left
 * out, so no B index is listed twice.
 * Outputs as annf_field_dev: nn_out is the first list entry, dist_out its
 * exact L2 (correctly rounded), knn_out the K list (-1 past the candidates).
 * No best-of-K pass: the list is already ranked by the full patch distance.
 * The tree this call builds keeps no points in tree order (the full-patch
 * path reads only its nodes, order and leaves): annf_neighbors,
 * annf_process_rows and annf_tree_export of xi_pts on the handle return
 * ANNF_ENOTREE until a tree is fitted again.
 * Stage times (annf_stage_times): 0 project B, 1 launch of A's projection,
 * 2 tree build (ends once A is projected too), 3 natural leaves +
 * bucketing, 4 MFMA leaf k-NN, 5 propagated pass + merge. */
#define ANNF_PREC_BF16 0
#define ANNF_PREC_FP8 1
int annf_field_mm_dev(annf_ctx *ctx, const uint8_t *img_a, int ha, int wa,
                      const uint8_t *img_b, int hb, int wb, int psize,
                      const void *basis, const double *mean, int d,
                      int projection, int tree_depth, int n_neighbors,
                      int nprobe, int precision, int shortlist, int propagate,
                      int *nn_out, float *dist_out, int *knn_out,
                      float *red_a_out, float *red_b_out);

/* The FP8 operands of annf_field_mm_dev's ANNF_PREC_FP8 form (no
 * counterpart in the reference; the codes its leaf pass reads): for every p x p
 * patch i of a device image (h x w x 3 uint8, psize <= 12), m = max |x - 128|
 * over the patch, scale s = 448 / m (1 when m = 0), and codes[i * DQ + r * RW
 * + k] the E4M3 codes (round to nearest even) of (x - 128) * s for bytes
 * 4k .. 4k + 3 of patch row r (RW = ceil(3p / 4) dwords per row, zero codes
 * past the row and up to DQ = 64 * ceil(p * RW * 4 / 64) bytes per patch);
 * inv[i] = m / 448 (1 when m = 0); norm[i] = (sum of the squared code values)
 * * inv[i]^2 (the sum exact, the product in FP64, rounded to FP32).  Any
 * output may be null. */
int annf_patch_fp8_dev(annf_ctx *ctx, const uint8_t *img, int h, int w,
                       int psize, uint8_t *codes, float *norm, float *inv);

/* Projection alone (openmp/base.py:193-200's transform, or the handle's
 * encoder): every p x p patch of a device image (h x w x 3 uint8) into out
 * (n x d float32, patch order of the reference's extract_patches_2d), with
 * the same arithmetic as annf_field_dev's projection of that image. */
int annf_project_dev(annf_ctx *ctx, const uint8_t *img, int h, int w,
                     int psize, const void *basis, const double *mean, int d,
                     int projection, float *out);

/* ---- the openclopt backend's semantics (replaces
 *      annfmp/openclopt/src/base.c fit_extern:257-520, which runs
 *      futhark/driverKNN.fut's buildKDtree, findNaturalLeavesFixK,
 *      exactKnnFixK, propagateFixK and selectBestNN) --------------------
 * On reduced patches (device float32, A: n_rows * n_cols x d, B: n_b x d):
 * the tree of buildKDtree.fut:19-144 (ceil(n_b / leaf_size) rounded to the
 * nearest power of two leaves of ppl = ceil(n_b / leaves) points, padded
 * with +inf rows; split dimension = the longest side of the box the
 * ancestors' medians cut from the global bounds; split value = the mean of
 * the sorted node's middle pair; stable sort), descent left on <=
 * (kd-traverse.fut:10), K fixed to 8 (driverKNN.fut:16): natural-leaf lists,
 * the exact 8-NN for row 0 (the stackless traverseOnce, one query per
 * lane), then row by row the leaves of the B patches one row below the row
 * above's list, without the natural leaf and repeats (driverKNN.fut:326-349).
 * Outputs (device, any may be NULL): knn_out (n_a x 8 original B indices,
 * -1 where a list is unfilled), knd_out (n_a x 8 squared reduced
 * distances), nn_out (n_a: the first strict minimum of the exact full-patch
 * distance over the 8, selectBestNN; needs both images, h x w x 3 uint8,
 * whose patch grids match the reduced sets).  n_rows > 1 needs n_b = n_a (the
 * row chain reads B indices on A's grid).  FP32 operations are rounded one
 * by one. */
int annf_field_openclopt_dev(annf_ctx *ctx, const float *red_a, int n_rows,
                             int n_cols, const float *red_b, int n_b, int d,
                             int leaf_size, const uint8_t *img_a, int ha,
                             int wa, const uint8_t *img_b, int hb, int wb,
                             int psize, int *nn_out, int *knn_out,
                             float *knd_out);
/* The last openclopt-semantics tree of the handle, to host buffers (any may
 * be NULL): shape = {height, internal nodes q, points per leaf, padded
 * count m'}, dims / vals / clanc (q: split dimension, median, closest
 * ancestor splitting the same dimension or -1), indir (m': tree position ->
 * point, >= n_b for pads), orig2leaf (m': point -> leaf). */
int annf_openclopt_tree_export(annf_ctx *ctx, int *shape, int *dims,
                               float *vals, int *clanc, int *indir,
                               int *orig2leaf);

/* ---- learned patch encoder (no reference counterpart; replaces the PCA
 *      projection of openmp/base.py:193-200 as the tree's feature map) ----
 * An MLP D -> h1 -> h2 -> d (D = p*p*3), GELU (tanh form) after the hidden
 * layers, BF16 operands on MFMA with FP32 accumulation.  A patch byte x
 * enters as (x - 128) / 128.  params is one FP32 device buffer:
 *   W1 (h1 x D), b1 (h1), W2 (h2 x h1), b2 (h2), W3 (d x h2), b3 (d),
 * rows of W the output units, so layer l computes W_l h + b_l.
 * x: device uint8 patch rows (n x D, sklearn extract_patches_2d order). */
int annf_mlp_forward_dev(annf_ctx *ctx, const uint8_t *x, int n, int D,
                         const float *params, int h1, int h2, int d,
                         float *out);
/* Triplet margin loss (PyTorch TripletMarginLoss(margin, p=2, eps=1e-6),
 * mean over the batch) of the B triplets whose patches are the rows of x:
 * B queries, then their B positives, then B negatives (3B x D).  Writes
 * the gradient of the loss w.r.t. params into grads (same layout) and, if
 * loss is not NULL, the loss (one device float). */
int annf_mlp_triplet_grad_dev(annf_ctx *ctx, const uint8_t *x, int B, int D,
                              const float *params, int h1, int h2, int d,
                              float margin, float *grads, float *loss);
/* The same gradient with no device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight training loops).  Its workspace is a fixed per-GPU array of 4 Mi
 * floats (16 MiB, held until the process ends): batches needing more
 * (3B x (2 h1 + 2 h2 + 2 d) + B + 1 floats; B <= 2568 at 128-128-16) return
 * ANNF_EUNSUPPORTED.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two handles, two threads) race. */
int annf_mlp_triplet_grad_noalloc_dev(annf_ctx *ctx, const uint8_t *x, int B,
                                      int D, const float *params, int h1,
                                      int h2, int d, float margin,
                                      float *grads, float *loss);
/* One Adam step (bias-corrected, step >= 1) on n parameters in place. */
int annf_mlp_adam_dev(annf_ctx *ctx, float *params, const float *grads,
                      float *m, float *v, long long n, float lr, float beta1,
                      float beta2, float eps, int step);
/* Installs params (copied, BF16) as the handle's encoder for patch size
 * psize; the fused kernel needs h1 = h2 = 128 and d <= 32. */
int annf_set_encoder(annf_ctx *ctx, const float *params, int psize, int h1,
                     int h2, int d);
/* Fused extraction + encoding of every p x p patch of a device image
 * (h x w x 3 uint8) into out (n x d float32, patch order of a3). */
int annf_encode_image_dev(annf_ctx *ctx, const uint8_t *img, int h, int w,
                          int psize, float *out);

/* ---- PCA covariance in FP64 (the covariance_eigh solver of scikit-learn
 *      1.7.2 that _fit_pca_model reaches, openmp/base.py:168-191) ---------
 * Mean (D doubles) and covariance (D x D doubles, row-major) of the patches
 * idx_a (n_a device int32 indices into img_a's patches) followed by idx_b
 * (n_b into img_b's), D = p*p*3: mean = sum / n; C = X^T X;
 * C -= (n * mean_col) * mean_row; C /= n - 1 -- sklearn's expression; the
 * sums are exact, so the result equals it bit for bit.  n_a + n_b <= 2^24. */
int annf_pca_cov_dev(annf_ctx *ctx, const uint8_t *img_a, int ha, int wa,
                     const int *idx_a, int n_a, const uint8_t *img_b, int hb,
                     int wb, const int *idx_b, int n_b, int psize,
                     double *mean_out, double *cov_out);
/* The whole PCA fit on the GPU: the covariance above, its eigendecomposition
 * in FP64 (rocSOLVER dsyevd, divide and conquer; scikit-learn calls LAPACK's
 * eigh), the d leading eigenvectors in descending eigenvalue order with
 * scikit-learn's sign rule (svd_flip, u_based_decision=False: the first
 * entry of largest magnitude of each component positive).  Writes
 * components_out (d x D doubles, row-major, PCA.components_), mean_out (D
 * doubles, PCA.mean_) and, when not NULL, variance_out (d doubles,
 * PCA.explained_variance_).  Equal to scikit-learn's fit up to the
 * eigensolver's rounding (a few ulps of the leading components). */
int annf_pca_fit_dev(annf_ctx *ctx, const uint8_t *img_a, int ha, int wa,
                     const int *idx_a, int n_a, const uint8_t *img_b, int hb,
                     int wb, const int *idx_b, int n_b, int psize, int d,
                     double *components_out, double *mean_out,
                     double *variance_out);

/* ---- exact field (evaluation reference; no reference counterpart) -----
 * For the nq A patches listed in query_idx (device int32 patch indices),
 * the k nearest B patches by exact squared L2 over the full p*p*3 uint8
 * patch, ordered by (distance, index).  Device pointers throughout; out_idx
 * and out_d2 are nq x k int32 (original B index, squared distance). */
int annf_exact_field_dev(annf_ctx *ctx, const uint8_t *img_a, int ha, int wa,
                         const uint8_t *img_b, int hb, int wb, int psize,
                         const int *query_idx, int nq, int k, int *out_idx,
                         int *out_d2);

/* Per-stage device times (ms) of the last annf_field_dev /
 * annf_process_rows / annf_field_mm_dev call, measured with HIP events on the
 * handle's main stream.  The query image A is projected on the handle's second
 * stream while the tree over B is built, so stage order is: 0 project B,
 * 1 launch of A's projection, 2 tree build (ends once A is projected too),
 * then for annf_field_dev 3 natural leaves + leaf scan, 4 row 0 exact search,
 * 5 row propagation sweep, 6 conversion + best-of-K, and for
 * annf_field_mm_dev 3 natural leaves + bucketing, 4 leaf k-NN, 5 propagated
 * pass + merge.  Returns the number of stages written (<= max_stages). */
int annf_stage_times(annf_ctx *ctx, float *ms, int max_stages);
/* Enables (1) or disables (0) the per-stage events. */
int annf_set_timing(annf_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif

#endif /* ANNF_HIP_H */
