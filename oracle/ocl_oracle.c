/*
 * ocl_oracle.c -- CPU restatement of the openclopt backend's semantics
 * (annfmp/openclopt/src/futhark/*.fut, driven by openclopt/src/base.c:257-520).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the library's openclopt-semantics
 * mode.  Nothing in annfmp_amd/ links or loads it.
 *
 * The reference runs this pipeline as Futhark-generated CUDA.  Neither the
 * Futhark compiler nor CUDA is available, and its sort comes from the
 * un-vendored diku-dk/sorts 0.3.14 (futhark.pkg), so this restatement follows
 * the Futhark sources and that library's published radix_sort_float (a
 * stable LSD radix sort; negative floats have every bit flipped, others the
 * sign bit, so -0 sorts before +0).  Where the sources leave a choice open
 * the restatement fixes one, the same as the GPU mode:
 *   - FP32 operations are rounded one by one (no fused multiply-add; NVRTC
 *     may contract some of the reference's multiply-adds);
 *   - the longest-width reduction (buildKDtree.fut:73-76, a reduce_comm whose
 *     operator keeps its left operand on ties) keeps the first dimension;
 *   - the best-of-8 full-patch distance (driverKNN.fut:93-125, f32.sum) is
 *     the exact integer sum, which is what any summation order gives below
 *     2^24;
 *   - an unfilled list entry (-1, possible only in a leaf with fewer than 8
 *     real points) is skipped where the reference would index with it.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OCL_K 8 /* driverKNN.fut:16, kk = 8 */

/* buildKDtree.fut:4-14: the nearest power-of-two exponent, ties down */
static int closest_log2(int p) {
    if (p <= 1) return 0;
    int r = 0;
    for (int q = p; q > 1; q >>= 1) ++r;
    const long long down = (long long)p - (1LL << r), up = (1LL << (r + 1)) - p;
    return down <= up ? r : r + 1;
}

/* buildKDtree.fut:19-30: (height without leaves, internal nodes, points per
 * leaf, padded point count) */
void ocl_tree_shape(int m, int defppl, int *height, int *q, int *ppl, int *mprime) {
    const int dnl = (m + defppl - 1) / defppl;
    const int hp1 = closest_log2(dnl);
    if (hp1 <= 0) {
        *height = -1;
        *q = 0;
        *ppl = m;
        *mprime = m;
        return;
    }
    const int h = hp1 - 1, nl = 1 << (h + 1);
    *height = h;
    *q = nl - 1;
    *ppl = (m + nl - 1) / nl;
    *mprime = nl * *ppl;
}

/* radix_sort_float's order (diku-dk/sorts 0.3.14) */
static uint32_t fkey(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* point idx of the padded set: rows past m are +inf (buildKDtree.fut:88-89) */
static float pval(const float *X, int m, int d, int idx, int j) {
    return idx < m ? X[(size_t)idx * d + j] : INFINITY;
}

typedef struct {
    uint32_t key;
    int pos;
    int idx;
} ocl_item;

static int item_cmp(const void *a, const void *b) {
    const ocl_item *x = (const ocl_item *)a, *y = (const ocl_item *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);   /* stable */
}

/* buildKDtree.fut:46-144 and 150-157.  Outputs: dims / vals / clanc [q],
 * indir [m'] (tree position -> padded point), leaf_pts [m' x d] (rows in
 * tree order, +inf for pads) and orig2leaf [m'] (point -> leaf). */
void ocl_build_tree(const float *X, int m, int d, int defppl, int *dims, float *vals, int *clanc, int *indir,
                    float *leaf_pts, int *orig2leaf) {
    int height, q, ppl, mp;
    ocl_tree_shape(m, defppl, &height, &q, &ppl, &mp);
    float *lubs = (float *)malloc(sizeof(float) * 2 * d), *cur = (float *)malloc(sizeof(float) * 2 * d);
    for (int j = 0; j < d; ++j) {
        float lo = FLT_MAX, hi = -FLT_MAX;
        for (int i = 0; i < m; ++i) {
            const float v = X[(size_t)i * d + j];
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
        lubs[j] = lo;
        lubs[d + j] = hi;
    }
    for (int i = 0; i < mp; ++i) indir[i] = i;
    ocl_item *items = (ocl_item *)malloc(sizeof(ocl_item) * (size_t)(mp > 0 ? mp : 1));
    for (int lev = 0; lev <= height; ++lev) {
        const int nodes = 1 << lev, ppn = mp / nodes;
        for (int i = 0; i < nodes; ++i) {
            const int node = i + nodes - 1;
            /* updateBounds (buildKDtree.fut:32-44): walking down from the
             * root, a right child's lower bound and a left child's upper
             * bound become the ancestor's median unless it is infinite */
            memcpy(cur, lubs, sizeof(float) * 2 * d);
            int anc = 0;
            for (int t = 0; t < lev; ++t) {
                const int k = lev - t - 1;
                const int child = (node + 1 - (1 << k)) / (1 << k);
                const int adim = dims[anc];
                const int li = (child & 1) == 0 ? adim : d + adim;
                if (!isinf(vals[anc])) cur[li] = vals[anc];
                anc = child;
            }
            /* the dimension of largest |ub - lb| (first on ties) */
            int best = -1;
            float bv = -FLT_MAX;
            for (int j = 0; j < d; ++j) {
                const float w = fabsf(cur[d + j] - cur[j]);
                if (!(bv >= w)) { best = j; bv = w; }
            }
            /* findClosestMed (buildKDtree.fut:46-55) */
            int c = node, res = -1;
            while (c != 0 && res == -1) {
                const int par = (c - 1) / 2;
                res = dims[par] == best ? par : -1;
                c = par;
            }
            dims[node] = best;
            clanc[node] = res;
            /* stable sort of the node's chunk by the chosen dimension */
            int *chunk = indir + (size_t)i * ppn;
            for (int t = 0; t < ppn; ++t) {
                items[t].key = fkey(pval(X, m, d, chunk[t], best));
                items[t].pos = t;
                items[t].idx = chunk[t];
            }
            qsort(items, (size_t)ppn, sizeof(ocl_item), item_cmp);
            for (int t = 0; t < ppn; ++t) chunk[t] = items[t].idx;
            const int mi = ppn / 2;
            const float a = pval(X, m, d, chunk[mi], best), b = pval(X, m, d, chunk[mi - 1], best);
            vals[node] = (a + b) / 2.0f;
        }
    }
    if (leaf_pts)
        for (int t = 0; t < mp; ++t)
            for (int j = 0; j < d; ++j) leaf_pts[(size_t)t * d + j] = pval(X, m, d, indir[t], j);
    if (orig2leaf)
        for (int t = 0; t < mp; ++t) orig2leaf[indir[t]] = t / ppl;
    free(items);
    free(cur);
    free(lubs);
}

/* kd-traverse.fut:5-14: left iff q[dim] <= median */
int ocl_find_leaf(const int *dims, const float *vals, int height, const float *qv) {
    const int q = (1 << (height + 1)) - 1;
    int node = 0;
    while (node < q) node = qv[dims[node]] <= vals[node] ? 2 * node + 1 : 2 * node + 2;
    return node - q;
}

/* util.fut:19-21 sumSqrsSeq: r = r + z*z, z = query - point, in order */
static float sqdist(const float *x, const float *y, int d) {
    float r = 0.0f;
    for (int j = 0; j < d; ++j) {
        const float z = x[j] - y[j];
        const float zz = z * z;
        r = r + zz;
    }
    return r;
}

/* brute-force.fut:5-22: sequential insert, an equal distance never
 * displaces; positions are beg + i */
void ocl_brute_force(const float *qv, int d, const float *pts, int beg, int npts, int *ki, float *kd) {
    for (int i = 0; i < npts; ++i) {
        float dist = sqdist(qv, pts + (size_t)i * d, d);
        if (dist > kd[OCL_K - 1]) continue;
        int ref = beg + i;
        for (int j = 0; j < OCL_K; ++j) {
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

/* brute-force.fut:25-66 (bruteForcePar + sortPartSortedSeqs): the leaf's
 * smallest (distance, position) pairs replace the list's largest entries
 * from the end while strictly smaller; the two sorted runs are then merged
 * taking the back run on ties */
void ocl_brute_force_par(const float *qv, int d, const float *pts, int beg, int npts, int *ki, float *kd) {
    float *dist = (float *)malloc(sizeof(float) * (size_t)(npts > 0 ? npts : 1));
    for (int i = 0; i < npts; ++i) dist[i] = sqdist(qv, pts + (size_t)i * d, d);
    for (int j = 0; j < OCL_K; ++j) {
        int bi = npts;
        float bv = INFINITY;
        for (int i = 0; i < npts; ++i) {
            if (dist[i] < bv) { bv = dist[i]; bi = i; }
            else if (dist[i] == bv && i < bi) bi = i;
        }
        if (!(bv < kd[OCL_K - 1 - j])) break;
        dist[bi] = FLT_MAX;
        ki[OCL_K - 1 - j] = beg + bi;
        kd[OCL_K - 1 - j] = bv;
    }
    int oi[OCL_K];
    float od[OCL_K];
    int b = 0, e = OCL_K - 1;
    for (int i = 0; i < OCL_K; ++i) {
        if (kd[b] < kd[e]) { oi[i] = ki[b]; od[i] = kd[b]; ++b; }
        else { oi[i] = ki[e]; od[i] = kd[e]; --e; }
    }
    memcpy(ki, oi, sizeof oi);
    memcpy(kd, od, sizeof od);
    free(dist);
}

static int bit_of(int stk, int i) { return (stk >> i) & 1; }
static int set_bit(int stk, int i, int v) { return v ? (stk | (1 << i)) : (stk & ~(1 << i)); }
static int level_of(int node) {
    int l = 0;
    for (int y = node + 1; y > 1; y >>= 1) ++l;
    return l;
}

/* kd-traverse.fut:44-57 getAncSameDimContrib: the first visited ancestor on
 * the chain of closest same-dimension ancestors gives median - q */
static float anc_contrib(const int *clanc, const float *vals, float qmi, int stack, int node) {
    int idx = node;
    while (idx >= 0) {
        const int anc = clanc[idx];
        if (anc == -1) return 0.0f;
        if (!bit_of(stack, level_of(anc))) idx = anc;
        else return vals[anc] - qmi;
    }
    return 0.0f;
}

/* kd-traverse.fut:27-116 traverseOnce: the next leaf to scan after `leaf`
 * (q + 1 when done); stack bit l: the level-l node's second child was
 * visited; dist: the incremental squared distance to the current cell */
void ocl_traverse_once(int height, const int *dims, const float *vals, const int *clanc, const float *qv, float wnn,
                       int *leaf, int *stack, float *dist) {
    const int q = (1 << (height + 1)) - 1;
    if (*leaf >= q + 1) return;
    int node = *leaf + q, count = height, rec = -1, stk = *stack;
    float ds = *dist;
    while (node != 0 && rec < 0) {
        const int par = (node - 1) / 2;
        const int scnd = bit_of(stk, count);
        const float qmd = qv[dims[par]];
        const float cur = vals[par] - qmd;
        const float cur2 = cur * cur;
        const float prv = anc_contrib(clanc, vals, qmd, stk, par);
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
            stk = set_bit(stk, count, 1);
            rec = (node % 2 == 0) ? node - 1 : node + 1;
            ds = dplus;
            node = par;
        }
    }
    if (node == 0 && rec == -1) {
        *leaf = q + 1;
        *stack = stk;
        *dist = ds;
        return;
    }
    int n = rec;
    while (n < q) {
        ++count;
        stk = set_bit(stk, count, 0);
        n = qv[dims[n]] <= vals[n] ? 2 * n + 1 : 2 * n + 2;
    }
    *leaf = n - q;
    *stack = stk;
    *dist = ds;
}

/* driverKNN.fut:209-243 exactKnn for one query, its list (ki, kd) already
 * holding the natural leaf's brute force */
void ocl_exact_knn(int height, const int *dims, const float *vals, const int *clanc, const float *leaf_pts, int d,
                   int ppl, const float *qv, int nat_leaf, int *ki, float *kd) {
    const int nleaves = 1 << (height + 1);
    int leaf = nat_leaf, stack = 0;
    float dist = 0.0f;
    for (;;) {
        ocl_traverse_once(height, dims, vals, clanc, qv, kd[OCL_K - 1], &leaf, &stack, &dist);
        if (leaf >= nleaves) break;
        ocl_brute_force_par(qv, d, leaf_pts + (size_t)leaf * ppl * d, leaf * ppl, ppl, ki, kd);
    }
}

/* driverKNN.fut:250-261 estimateIndex: the leaf of the B patch one row
 * below (clamped to the last row) the point at tree position pos */
static int estimate_leaf(const int *indir, const int *orig2leaf, int n_rows, int n_cols, int pos) {
    const int o = indir[pos];
    const int y = o / n_cols, x = o - y * n_cols;
    const int cy = y < n_rows - 1 ? y + 1 : n_rows - 1;
    return orig2leaf[cy * n_cols + x];
}

/* The whole openclopt pipeline on reduced patches (base.c:257-520):
 * red_a n_a x d (n_a = n_rows * n_cols), red_b m x d; defppl = leaf_size.
 * Outputs: knn [n_a x 8] original B indices, knd [n_a x 8] squared reduced
 * distances, nat [n_a] natural leaves.  best_of (when pa / pb are given,
 * D bytes per patch): nn [n_a] = first strict minimum of the exact full-patch
 * distance over the 8 candidates (driverKNN.fut:93-125).  Also returns the
 * tree (dims, vals, clanc [q], indir, orig2leaf [m']).  rows_limit > 0
 * computes only the first rows_limit rows (each row depends only on the
 * rows above it); the other rows' outputs are left as they are. */
void ocl_field(const float *red_a, int n_rows, int n_cols, const float *red_b, int m, int d, int defppl,
               const uint8_t *pa, const uint8_t *pb, int D, int *nn, int *knn, float *knd, int *nat, int *dims,
               float *vals, int *clanc, int *indir, int *orig2leaf, int rows_limit) {
    int height, q, ppl, mp;
    ocl_tree_shape(m, defppl, &height, &q, &ppl, &mp);
    float *lp = (float *)malloc(sizeof(float) * (size_t)mp * d);
    ocl_build_tree(red_b, m, d, defppl, dims, vals, clanc, indir, lp, orig2leaf);
    const int nr = rows_limit > 0 && rows_limit < n_rows ? rows_limit : n_rows;
    const int n_a = nr * n_cols;
    int *pos = (int *)malloc(sizeof(int) * (size_t)n_a * OCL_K);
    /* findNaturalLeaves (driverKNN.fut:133-152) */
    for (int i = 0; i < n_a; ++i) {
        const float *qv = red_a + (size_t)i * d;
        const int leaf = ocl_find_leaf(dims, vals, height, qv);
        nat[i] = leaf;
        int *ki = pos + (size_t)i * OCL_K;
        float *kd = knd + (size_t)i * OCL_K;
        for (int k = 0; k < OCL_K; ++k) { ki[k] = -1; kd[k] = FLT_MAX; }
        ocl_brute_force(qv, d, lp + (size_t)leaf * ppl * d, leaf * ppl, ppl, ki, kd);
    }
    /* row 0: exact k-NN (exactKnnFixK with s = n_cols) */
    for (int x = 0; x < n_cols && x < n_a; ++x)
        ocl_exact_knn(height, dims, vals, clanc, lp, d, ppl, red_a + (size_t)x * d, nat[x], pos + (size_t)x * OCL_K,
                      knd + (size_t)x * OCL_K);
    /* rows 1..: the leaves of the B patches below the row above's list
     * (driverKNN.fut:287-350), without the natural leaf and repeats, in
     * list order, each merged by bruteForcePar into the natural-leaf list */
    for (int y = 1; y < nr; ++y)
        for (int x = 0; x < n_cols; ++x) {
            const int i = y * n_cols + x;
            const int *up = pos + (size_t)(i - n_cols) * OCL_K;
            int ul[OCL_K], nu = 0;
            for (int k = 0; k < OCL_K; ++k) {
                if (up[k] < 0) continue;
                const int l = estimate_leaf(indir, orig2leaf, n_rows, n_cols, up[k]);
                int found = l == nat[i];
                for (int j = 0; j < nu && !found; ++j) found = ul[j] == l;
                if (!found) ul[nu++] = l;
            }
            const float *qv = red_a + (size_t)i * d;
            for (int j = 0; j < nu; ++j)
                ocl_brute_force_par(qv, d, lp + (size_t)ul[j] * ppl * d, ul[j] * ppl, ppl, pos + (size_t)i * OCL_K,
                                    knd + (size_t)i * OCL_K);
        }
    /* positions -> original indices (driverKNN.fut:409) */
    for (size_t t = 0; t < (size_t)n_a * OCL_K; ++t) knn[t] = pos[t] >= 0 ? indir[pos[t]] : -1;
    if (pa && pb && nn)
        for (int i = 0; i < n_a; ++i) {
            int best = -1;
            long long bd = -1;
            const uint8_t *a = pa + (size_t)i * D;
            for (int k = 0; k < OCL_K; ++k) {
                const int j = knn[(size_t)i * OCL_K + k];
                if (j < 0) continue;
                const uint8_t *b = pb + (size_t)j * D;
                long long s = 0;
                for (int f = 0; f < D; ++f) {
                    const int t = (int)b[f] - (int)a[f];
                    s += (long long)t * t;
                }
                if (best < 0 || s < bd) { best = j; bd = s; }
            }
            nn[i] = best;
        }
    free(pos);
    free(lp);
}
