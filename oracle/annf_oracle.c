/*
 * annf_oracle.c -- CPU restatement of annfmp's openmp k-d-tree NNF backend.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * annfmp_amd/csrc and the timed CPU baseline in bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never does.
 *
 * It restates, from the reference's behaviour (not its code), the FP32
 * semantics of:
 *   - tree build           annfmp/openmp/kdtree/src/kdtree.c:76-117, 363-426, 463-489
 *                          helper.c:63-113, util.c:159-188
 *   - natural-leaf descent kdtree.c:171-214 (only_leaf_idx=1, max_leaves=1)
 *   - exact k-NN search    kdtree.c:129-268 (row 0, neighbors_extern)
 *   - leaf scan / insert   kdtree.c:286-352, 502-532; helper.c:164-179
 *   - row driver           base.c:150-360 (row 0, rows >= 1, propagation,
 *                          index conversion, best-of-K)
 *   - norm_array           base.c:362-374
 *
 * Floating point: compiled with -ffp-contract=off and without fast-math, so
 * every distance is the sequential FP32 sum sum_j (p_j - q_j)^2 with one
 * rounding per operation, as in the reference built with -std=c99 -O3 on
 * baseline x86-64 (no FMA).
 *
 * Unlike the reference, no state is shared between OpenMP threads (the
 * reference's shared loop temporaries at base.c:181-182 are not reproduced),
 * every routine returns a status instead of calling exit(), and the
 * partition uses the closed form below instead of a copy-and-swap loop.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_FLT_MAX 3.402823466e+38f /* initial list distance, float.h:24 */

typedef struct {
    int n, d, depth;
    float *pts;    /* n*d, XI order (tree positions)            */
    int *orig;     /* n, original index of each tree position   */
    int *axis;     /* 2^depth - 1                               */
    float *split;  /* 2^depth - 1                               */
    int *leaves;   /* 2 * 2^depth: [from, to) per leaf          */
    int *pli;      /* n, leaf of each ORIGINAL index            */
} orc_tree;

/* ------------------------------------------------------------------ */
/* distance and k-list                                                 */
/* ------------------------------------------------------------------ */

/* helper.c:164-179: sequential FP32 sum of (p_j - q_j)^2. */
static float orc_dist(const float *p, const float *q, int d)
{
    float s = 0.0f;
    for (int j = 0; j < d; ++j) {
        float t = p[j] - q[j];
        s += t * t;
    }
    return s;
}

/* kdtree.c:502-532: keep the K best of the candidate stream, ties broken by
 * arrival order.  A candidate is dropped when the current K-th distance is
 * <= its distance; otherwise it takes the last slot and moves forward while
 * strictly smaller than its predecessor. */
static void orc_insert(float dist, int pos, float *ld, int *lp, int K)
{
    if (ld[K - 1] <= dist) return;
    int j = K - 1;
    ld[j] = dist;
    lp[j] = pos;
    while (j > 0 && ld[j] < ld[j - 1]) {
        float td = ld[j]; ld[j] = ld[j - 1]; ld[j - 1] = td;
        int tp = lp[j]; lp[j] = lp[j - 1]; lp[j - 1] = tp;
        --j;
    }
}

static void orc_list_reset(float *ld, int *lp, int K)
{
    for (int k = 0; k < K; ++k) { ld[k] = ORC_FLT_MAX; lp[k] = 0; }
}

/* kdtree.c:286-303: scan tree positions [from, to) in order. */
static void orc_scan_range(const orc_tree *t, const float *q, int from, int to,
                           float *ld, int *lp, int K)
{
    for (int i = from; i < to; ++i)
        orc_insert(orc_dist(t->pts + (size_t)i * t->d, q, t->d), i, ld, lp, K);
}

/* ------------------------------------------------------------------ */
/* tree build                                                          */
/* ------------------------------------------------------------------ */

/* Rank-k value (0-based) of v[0..n) under float '<', by quickselect on a
 * scratch copy.  Any correct selection returns a value of the same '<'
 * equivalence class as util.c:146-188 (Wirth). */
static float orc_select(float *v, int n, int k)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        float piv = v[lo + (hi - lo) / 2];
        int i = lo, j = hi;
        while (i <= j) {
            while (v[i] < piv) ++i;
            while (piv < v[j]) --j;
            if (i <= j) { float tt = v[i]; v[i] = v[j]; v[j] = tt; ++i; --j; }
        }
        if (k <= j) hi = j;
        else if (k >= i) lo = i;
        else return v[k];
    }
    return v[k];
}

/* Destination slot (0-based within the node) of every element of a node
 * segment of c elements, for split value v on axis a.  Closed form of the
 * two passes of helper.c:63-113:
 *   pass 1: x < v elements go to the front in segment order; the others
 *           fill the back from slot c-1 downwards (so block G = [L, c) holds
 *           them in reverse segment order);
 *   pass 2: for t = 0..E-1, slot L+t is exchanged with slot L+e_t, where
 *           e_t is the G-offset of the t-th element of G equal to v.
 * A '>' element starting at G-offset g is moved by step t = g (to e_g), then
 * possibly again by step t = e_g, ...; it ends at the first offset >= E of
 * the chain g, e_g, e_{e_g}, ...  Equal elements end at L+t in G order. */
static void orc_partition_slots(const float *col, int c, float v, int *slot, int *eq_off)
{
    int L = 0, E = 0;
    for (int i = 0; i < c; ++i) {
        if (col[i] < v) ++L;
        else if (col[i] == v) ++E;
    }
    int g = c - L;
    /* walk the segment in order; G offsets count down from g-1, and the
     * t-th equal element of G is the (E-1-t)-th one of the segment */
    int r_lt = 0, r_geq = 0, r_eq = 0;
    for (int i = 0; i < c; ++i) {
        if (col[i] < v) {
            slot[i] = r_lt++;
        } else {
            int goff = (g - 1) - r_geq++;
            if (col[i] == v) {
                int t = (E - 1) - r_eq++;
                eq_off[t] = goff;
                slot[i] = L + t;
            } else {
                slot[i] = -1 - goff; /* resolved below */
            }
        }
    }
    for (int i = 0; i < c; ++i) {
        if (slot[i] < 0) {
            int p = -1 - slot[i];
            while (p < E) p = eq_off[p];
            slot[i] = L + p;
        }
    }
}

/* kdtree.c:363-426, longest-box rule: per-dimension data box over the
 * segment, width = max - min in FP32, first strictly longest axis, initial
 * longest 0.0 (so all-zero widths give axis 0). */
static int orc_axis(const float *pts, int lo, int hi, int d, float *mn, float *mx)
{
    for (int j = 0; j < d; ++j) { mn[j] = ORC_FLT_MAX; mx[j] = -ORC_FLT_MAX; }
    for (int i = lo; i < hi; ++i) {
        const float *p = pts + (size_t)i * d;
        for (int j = 0; j < d; ++j) {
            if (p[j] < mn[j]) mn[j] = p[j];
            if (p[j] > mx[j]) mx[j] = p[j];
        }
    }
    float longest = 0.0f;
    int ax = 0;
    for (int j = 0; j < d; ++j) {
        float w = mx[j] - mn[j];
        if (w > longest) { longest = w; ax = j; }
    }
    return ax;
}

/*
 * Build the tree over X (n x d, row-major).  Level by level, each node is
 * the contiguous segment it owns in tree-position order (kdtree.c:76-117
 * visits the same segments depth-first; the result does not depend on the
 * visiting order).  Returns 0, or -1 on a bad argument / allocation.
 */
int orc_build_tree(const float *X, int n, int d, int depth, orc_tree *t)
{
    if (n < 1 || d < 1 || depth < 0 || depth > 30) return -1;
    if ((n >> depth) < 1) return -1; /* every node must be non-empty */
    t->n = n; t->d = d; t->depth = depth;
    memcpy(t->pts, X, (size_t)n * d * sizeof(float));
    for (int i = 0; i < n; ++i) t->orig[i] = i;

    float *tmp = (float *)malloc((size_t)n * d * sizeof(float));
    int *tmpo = (int *)malloc((size_t)n * sizeof(int));
    float *col = (float *)malloc((size_t)n * sizeof(float));
    float *sel = (float *)malloc((size_t)n * sizeof(float));
    int *slot = (int *)malloc((size_t)n * sizeof(int));
    int *eqo = (int *)malloc((size_t)n * sizeof(int));
    float *mn = (float *)malloc((size_t)d * sizeof(float));
    float *mx = (float *)malloc((size_t)d * sizeof(float));
    int *bnd = (int *)malloc(((size_t)(1u << depth) + 1) * sizeof(int));
    int *nbnd = (int *)malloc(((size_t)(1u << depth) + 1) * sizeof(int));
    if (!tmp || !tmpo || !col || !sel || !slot || !eqo || !mn || !mx || !bnd || !nbnd) {
        free(tmp); free(tmpo); free(col); free(sel); free(slot); free(eqo);
        free(mn); free(mx); free(bnd); free(nbnd);
        return -1;
    }
    bnd[0] = 0; bnd[1] = n;
    for (int lvl = 0; lvl < depth; ++lvl) {
        int nodes = 1 << lvl;
        for (int j = 0; j < nodes; ++j) {
            int lo = bnd[j], hi = bnd[j + 1], c = hi - lo;
            int node = nodes - 1 + j;
            int ax = orc_axis(t->pts, lo, hi, d, mn, mx);
            for (int i = 0; i < c; ++i) col[i] = t->pts[(size_t)(lo + i) * d + ax];
            memcpy(sel, col, (size_t)c * sizeof(float));
            float v = orc_select(sel, c, c / 2);
            orc_partition_slots(col, c, v, slot, eqo);
            for (int i = 0; i < c; ++i) {
                int s = lo + slot[i];
                memcpy(tmp + (size_t)s * d, t->pts + (size_t)(lo + i) * d, (size_t)d * sizeof(float));
                tmpo[s] = t->orig[lo + i];
            }
            memcpy(t->pts + (size_t)lo * d, tmp + (size_t)lo * d, (size_t)c * d * sizeof(float));
            memcpy(t->orig + lo, tmpo + lo, (size_t)c * sizeof(int));
            t->axis[node] = ax;
            t->split[node] = t->pts[(size_t)(lo + c / 2) * d + ax];
            nbnd[2 * j] = lo;
            nbnd[2 * j + 1] = lo + c / 2;
        }
        nbnd[2 * nodes] = n;
        int *sw = bnd; bnd = nbnd; nbnd = sw;
    }
    int nleaves = 1 << depth;
    for (int l = 0; l < nleaves; ++l) {
        t->leaves[2 * l] = bnd[l];
        t->leaves[2 * l + 1] = bnd[l + 1];
        for (int i = bnd[l]; i < bnd[l + 1]; ++i) t->pli[t->orig[i]] = l;
    }
    free(tmp); free(tmpo); free(col); free(sel); free(slot); free(eqo);
    free(mn); free(mx); free(bnd); free(nbnd);
    return 0;
}

/* ------------------------------------------------------------------ */
/* queries                                                             */
/* ------------------------------------------------------------------ */

/* kdtree.c:198-214 with only_leaf_idx=1: right iff q[axis] - split >= 0. */
int orc_natural_leaf(const orc_tree *t, const float *q)
{
    int idx = 0;
    for (int lvl = 0; lvl < t->depth; ++lvl) {
        float diff = q[t->axis[idx]] - t->split[idx];
        idx = (diff >= 0.0f) ? 2 * idx + 2 : 2 * idx + 1;
    }
    return idx - ((1 << t->depth) - 1);
}

typedef struct {
    const orc_tree *t;
    const float *q;
    float *ld;
    int *lp;
    int K;
    long visited;
    long max_leaves;
    int stop;
} orc_search;

/* kdtree.c:129-268: depth-first, near child first; the far child is
 * entered iff (q[axis]-split)^2 <= current K-th distance, evaluated after
 * the near subtree.  The search ends when more than max_leaves leaves would
 * be visited. */
static void orc_search_node(orc_search *s, int idx, int lvl)
{
    const orc_tree *t = s->t;
    if (s->stop) return;
    if (lvl == t->depth) {
        if (++s->visited > s->max_leaves) { s->stop = 1; return; }
        int leaf = idx - ((1 << t->depth) - 1);
        orc_scan_range(t, s->q, t->leaves[2 * leaf], t->leaves[2 * leaf + 1], s->ld, s->lp, s->K);
        return;
    }
    float diff = s->q[t->axis[idx]] - t->split[idx];
    int near_c = diff >= 0.0f ? 2 * idx + 2 : 2 * idx + 1;
    int far_c = diff >= 0.0f ? 2 * idx + 1 : 2 * idx + 2;
    orc_search_node(s, near_c, lvl + 1);
    if (s->stop) return;
    if (diff * diff <= s->ld[s->K - 1]) orc_search_node(s, far_c, lvl + 1);
}

void orc_exact_knn(const orc_tree *t, const float *q, int K, long max_leaves, float *ld, int *lp)
{
    orc_search s = {t, q, ld, lp, K, 0, max_leaves, 0};
    orc_list_reset(ld, lp, K);
    orc_search_node(&s, 0, 0);
}

/* kdtree.c:305-352: fresh list, scan the given leaves in order (a leaf
 * listed twice is scanned twice). */
void orc_scan_leaves(const orc_tree *t, const float *q, const int *leaf_list, int nl,
                     int K, float *ld, int *lp)
{
    orc_list_reset(ld, lp, K);
    for (int i = 0; i < nl; ++i) {
        int l = leaf_list[i];
        orc_scan_range(t, q, t->leaves[2 * l], t->leaves[2 * l + 1], ld, lp, K);
    }
}

/* neighbors_extern (base.c:88-132): exact k-NN, then original indices and
 * sqrt distances (sqrt taken in double, stored as float). */
int orc_neighbors(const orc_tree *t, const float *Q, int nq, int K, long max_leaves,
                  float *dist, int *idx, int nthreads)
{
    if (K < 1) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < nq; ++i)
        orc_exact_knn(t, Q + (size_t)i * t->d, K, max_leaves, dist + (size_t)i * K, idx + (size_t)i * K);
    for (long i = 0; i < (long)nq * K; ++i) {
        idx[i] = t->orig[idx[i]];
        dist[i] = (float)sqrt((double)dist[i]);
    }
    return 0;
}

/* base.c:362-374: sqrt of the sequential FP32 sum of squared u8 differences. */
float orc_norm_array(const uint8_t *a, const uint8_t *b, int dim)
{
    float s = 0.0f;
    for (int i = 0; i < dim; ++i) {
        float df = (float)a[i] - (float)b[i];
        s += df * df;
    }
    return (float)sqrt((double)s);
}

/*
 * process_rows_extern (base.c:150-360).
 *   red       : n_a x d reduced query patches (A), row-major patch order
 *   pa, pb    : n_a x D and n_b x D uint8 patches (only read if select_best)
 *   final     : n_a output indices (original B patch indices)
 *   lists_pos : optional n_a x K output, k-lists in tree positions
 *   lists_dst : optional n_a x K output, their reduced-space distances
 * Rows start from an exact search (row 0); each later row scans the
 * natural leaf and, with propagate, the leaves holding "one row below" of
 * each of the K matches of the patch above (base.c:255-282).
 */
int orc_process_rows(const orc_tree *t, const uint8_t *pa, const uint8_t *pb, int D,
                     const float *red, int n_a, int n_rows, int n_cols, int K,
                     int propagate, int select_best, int *final,
                     int *lists_pos, float *lists_dst, int nthreads)
{
    if (K < 1 || n_rows < 1 || n_cols < 1 || (long)n_rows * n_cols != n_a) return -1;
    int d = t->d;
    int *pos = lists_pos ? lists_pos : (int *)malloc((size_t)n_a * K * sizeof(int));
    float *dst = lists_dst ? lists_dst : (float *)malloc((size_t)n_a * K * sizeof(float));
    if (!pos || !dst) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const long unlimited = 2147483640L; /* kdtree/base.py:127 */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int x = 0; x < n_cols; ++x)
        orc_exact_knn(t, red + (size_t)x * d, K, unlimited, dst + (size_t)x * K, pos + (size_t)x * K);

    for (int y = 1; y < n_rows; ++y) {
#ifdef _OPENMP
#pragma omp parallel
#endif
        {
            int ll[1 + 64];
            int *lleaf = (K + 1 <= 65) ? ll : (int *)malloc((size_t)(K + 1) * sizeof(int));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (int x = 0; x < n_cols; ++x) {
                long qi = (long)y * n_cols + x;
                const float *q = red + (size_t)qi * d;
                int nl = 1;
                lleaf[0] = orc_natural_leaf(t, q);
                if (propagate) {
                    const int *above = pos + (size_t)(qi - n_cols) * K;
                    for (int k = 0; k < K; ++k) {
                        int o = t->orig[above[k]];
                        if (o + n_cols < n_a) o += n_cols;
                        lleaf[1 + k] = t->pli[o];
                    }
                    nl = 1 + K;
                }
                orc_scan_leaves(t, q, lleaf, nl, K, dst + (size_t)qi * K, pos + (size_t)qi * K);
            }
            if (lleaf != ll) free(lleaf);
        }
    }

#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long i = 0; i < n_a; ++i) {
        const int *lp = pos + (size_t)i * K;
        if (select_best) {
            int best = -1;
            float bd = ORC_FLT_MAX;
            for (int k = 0; k < K; ++k) {
                int j = t->orig[lp[k]];
                float dd = orc_norm_array(pa + (size_t)i * D, pb + (size_t)j * D, D);
                if (dd < bd) { bd = dd; best = j; }
            }
            final[i] = best;
        } else {
            final[i] = t->orig[lp[0]];
        }
    }
    if (!lists_pos) free(pos);
    if (!lists_dst) free(dst);
    return 0;
}

/* Rows >= row_begin of process_rows only, for a bounded timing sample: the
 * lists of row row_begin-1 must already be in lists_pos. */
int orc_rows_range(const orc_tree *t, const float *red, int n_a, int n_cols, int K,
                   int propagate, int row_begin, int row_end, int *lists_pos,
                   float *lists_dst, int nthreads)
{
    int d = t->d;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    for (int y = row_begin; y < row_end; ++y) {
#ifdef _OPENMP
#pragma omp parallel
#endif
        {
            int lleaf[1 + 64];
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (int x = 0; x < n_cols; ++x) {
                long qi = (long)y * n_cols + x;
                const float *q = red + (size_t)qi * d;
                int nl = 1;
                lleaf[0] = orc_natural_leaf(t, q);
                if (propagate && y > 0) {
                    const int *above = lists_pos + (size_t)(qi - n_cols) * K;
                    for (int k = 0; k < K && k < 64; ++k) {
                        int o = t->orig[above[k]];
                        if (o + n_cols < n_a) o += n_cols;
                        lleaf[1 + k] = t->pli[o];
                    }
                    nl = 1 + K;
                }
                orc_scan_leaves(t, q, lleaf, nl, K, lists_dst + (size_t)qi * K, lists_pos + (size_t)qi * K);
            }
        }
    }
    return 0;
}

/* Best-of-K straight from the two images (window reads), n_rows x n_cols
 * queries with K original indices each; same arithmetic as norm_array on
 * the patch f = (i*p + j)*3 + ch (openmp/base.py:152-166). */
int orc_best_of_k_images(const uint8_t *ia, const uint8_t *ib, int W, int p,
                         int n_rows, int n_cols, const int *cand, int K,
                         int *final, float *fdist, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long i = 0; i < (long)n_rows * n_cols; ++i) {
        int ya = (int)(i / n_cols), xa = (int)(i % n_cols);
        int best = -1;
        float bd = ORC_FLT_MAX;
        for (int k = 0; k < K; ++k) {
            int j = cand[(size_t)i * K + k];
            int yb = j / n_cols, xb = j % n_cols;
            float s = 0.0f;
            for (int r = 0; r < p; ++r) {
                const uint8_t *ra = ia + ((size_t)(ya + r) * W + xa) * 3;
                const uint8_t *rb = ib + ((size_t)(yb + r) * W + xb) * 3;
                for (int b = 0; b < 3 * p; ++b) {
                    float df = (float)ra[b] - (float)rb[b];
                    s += df * df;
                }
            }
            float dd = (float)sqrt((double)s);
            if (dd < bd) { bd = dd; best = j; }
        }
        final[i] = best;
        if (fdist) fdist[i] = bd;
    }
    return 0;
}
