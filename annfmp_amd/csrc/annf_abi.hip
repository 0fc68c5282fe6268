// C ABI of libannf_hip.so (declared in include/annf_hip.h): handles, device
// workspaces and the host-side sequencing of the kernels in annf_kernels.hip.
#include "annf_internal.h"
#include "../../include/annf_hip.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

using namespace annf;

// Workspace of annf_mlp_triplet_grad_noalloc_dev, the entry point that
// allocates nothing: one copy per GPU, held until the process ends.  Host
// code never passes `scratch` itself to a kernel (on the host it names a
// stand-in, not the device copy): it uses the device address
// hipGetSymbolAddress returns for the current device (scratch_address).
constexpr size_t kNoAllocFloats = (size_t)4 << 20;   // 16 MiB
static __device__ __align__(256) float scratch[kNoAllocFloats];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ANNF_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                                \
    } while (0)

#define RCHK(expr)                 \
    do {                           \
        int r_ = (expr);           \
        if (r_ != ANNF_OK) return r_; \
    } while (0)

// A growable device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return ANNF_OK;
        if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
        size_t want = bytes ? bytes : 16;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(ANNF_ENOMEM, "hipMalloc(%zu) failed", want);
        }
        cap = want;
        return ANNF_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

constexpr int kStages = 7;

}  // namespace

struct annf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // the query image's projection runs here, beside the tree build on `stream`
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // init_extern parameters (global.h:18-27)
    bool inited = false;
    int n_neighbors = 10, tree_depth = 0, max_leaves = 2147483640, num_threads = 1;
    int splitting_type = ANNF_SPLIT_LONGEST, verbosity = 0;
    // tree
    bool has_tree = false;
    TreeView tree;
    // tree storage
    DevBuf pts_a, pts_b, orig_a, orig_b, axis, split, leaves, pli;
    // build scratch
    DevBuf keys, eqoff, box, box2, nax, prefix, rank, hist, cnt, cnt_off, tot, bnd, errflag;
    // query workspaces
    DevBuf qbuf, natleaf, natdesc, bcount, boffset, bcursor, bchunk, order, p1d, p1p, p1l, propleaf;
    DevBuf fl_pos, fl_d, fl_l, final_i, final_d, stage_in, stage_in2, stage_in3;
    DevBuf basis32, mean32, basis64, mean64;
    DevBuf ev_am, ev_an, ev_bm, ev_bn, ev_part;
    // learned encoder: training workspace and the packed inference weights
    DevBuf mlp_ws, enc_w1, enc_b1, enc_w2, enc_b2, enc_w3, enc_b3;
    // full-patch leaf k-NN: per-query and per-pair exact keys, pair leaves
    DevBuf mm_keys1, mm_keys2, mm_pairleaf, mm_fkeys;
    DevBuf pca_scratch, pca_eig;
    // the openclopt-semantics mode: its tree and row lists
    DevBuf ocl_lub, ocl_indir, ocl_indir2, ocl_keys, ocl_keys2, ocl_dims, ocl_clanc, ocl_vals, ocl_leafpts,
        ocl_o2l, ocl_sort, ocl_nat, ocl_pos, ocl_dst;
    OclTree ocl;
    bool has_ocl = false;
    // rocSOLVER (the FP64 eigensolve of annf_pca_fit_dev), created on first use
    rocblas_handle blas = nullptr;
    DevBuf sort_keys, sort_tmp, mm_blkleaf, mm_ptorder, proj_pack;
    // exact BF16 leaf k-NN: centred BF16 images, per-patch norms
    DevBuf cimg_a, cimg_b, pnorm_a, pnorm_b, pnorm_tmp;
    // FP8 leaf k-NN: per-patch codes, norms and inverse scales of both
    // images, the per-(slot, leaf) approximate lists
    DevBuf f8_qa, f8_qb, f8_na, f8_nb, f8_ia, f8_ib, f8_keys;
    bool has_encoder = false;
    int enc_p = 0, enc_d = 0;
    // this device's copy of the allocation-free entry point's workspace
    float *noalloc_ws = nullptr;
    // timing
    bool timing = false;
    hipEvent_t ev[kStages + 1] = {};
    bool ev_used[kStages + 1] = {};
    float stage_ms[kStages] = {};
    bool have_stages = false;

    void mark(int i) {
        if (!timing) return;
        if (hipEventRecord(ev[i], stream) == hipSuccess) ev_used[i] = true;
    }
    void reset_marks() {
        for (int i = 0; i <= kStages; ++i) ev_used[i] = false;
        have_stages = false;
    }
    void collect_marks() {
        if (!timing) return;
        for (int i = 0; i < kStages; ++i) {
            stage_ms[i] = 0.0f;
            float ms = 0.0f;
            if (ev_used[i] && ev_used[i + 1] && hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess)
                stage_ms[i] = ms;
        }
        have_stages = true;
    }
    // host copies of the level bounds (kept alive for the async upload), and
    // the (n, depth) they and the device copy in bnd were computed for
    std::vector<int> h_bnd, h_off, h_cpn, h_maxc;
    long long bnd_key = -1;
    int bnd_max_leaf = 0;
    int *h_flag = nullptr;   // pinned host copy of the build's error flag

    void release_all() {
        DevBuf *list[] = {&pts_a, &pts_b, &orig_a, &orig_b, &axis, &split, &leaves, &pli, &keys, &eqoff, &box, &box2,
                          &nax, &prefix, &rank, &hist, &cnt, &cnt_off, &tot, &bnd, &errflag, &qbuf, &natleaf, &natdesc, &bcount, &boffset,
                          &bcursor, &bchunk, &order, &p1d, &p1p, &p1l, &propleaf, &fl_pos, &fl_d, &fl_l, &final_i,
                          &final_d, &stage_in, &stage_in2, &stage_in3, &basis32, &mean32, &basis64, &mean64,
                          &ev_am, &ev_an, &ev_bm, &ev_bn, &ev_part, &mlp_ws, &enc_w1, &enc_b1, &enc_w2,
                          &enc_b2, &enc_w3, &enc_b3, &mm_keys1, &mm_keys2, &mm_pairleaf, &mm_fkeys, &pca_scratch, &pca_eig, &ocl_lub, &ocl_indir, &ocl_indir2, &ocl_keys,
                          &ocl_keys2, &ocl_dims, &ocl_clanc, &ocl_vals, &ocl_leafpts, &ocl_o2l, &ocl_sort, &ocl_nat,
                          &ocl_pos, &ocl_dst,
                          &sort_keys, &sort_tmp, &mm_blkleaf, &mm_ptorder, &proj_pack, &cimg_a, &cimg_b,
                          &pnorm_a, &pnorm_b, &pnorm_tmp, &f8_qa, &f8_qb, &f8_na, &f8_nb, &f8_ia, &f8_ib,
                          &f8_keys};
        for (DevBuf *b : list) b->release();
        bnd_key = -1;
        has_tree = false;
        has_encoder = false;
        has_ocl = false;
    }
};

namespace {

int sync_stream(annf_ctx *c) {
    HIPCHK(hipStreamSynchronize(c->stream));
    return ANNF_OK;
}

// Waits for the stream and reads the device error flag of the last build;
// the flag's copy to pinned memory is queued behind the work, so the wait
// brings it back too.
int sync_and_check_build(annf_ctx *c) {
    int flag = 0;
    if (c->errflag.p && c->h_flag) {
        *c->h_flag = 0;
        HIPCHK(hipMemcpyAsync(c->h_flag, c->errflag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        RCHK(sync_stream(c));
        flag = *c->h_flag;
    } else {
        RCHK(sync_stream(c));
        if (c->errflag.p) HIPCHK(hipMemcpy(&flag, c->errflag.p, sizeof(int), hipMemcpyDeviceToHost));
    }
    if (flag) {
        c->has_tree = false;
        return fail(ANNF_EHIP, "tree build detected an inconsistent partition (flag %d)", flag);
    }
    return ANNF_OK;
}

int set_device(annf_ctx *c) {
    HIPCHK(hipSetDevice(c->device));
    return ANNF_OK;
}

// The device address of `scratch` on `device` (which must be current),
// looked up once per device and kept.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_addr[kMaxDevices] = {};

int scratch_address(int device, float **out) {
    if (device < 0 || device >= kMaxDevices) return fail(ANNF_EUNSUPPORTED, "device %d above %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(ANNF_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                        hipGetErrorString(e));
        g_scratch_addr[device] = static_cast<float *>(ptr);
    }
    *out = g_scratch_addr[device];
    return ANNF_OK;
}

// Segment bounds of every level (kdtree.c:114-115: left child gets c/2).
void level_bounds(int n, int depth, std::vector<int> &bnd, std::vector<int> &off, std::vector<int> &cpn,
                  std::vector<int> &maxc_out, int &max_leaf) {
    bnd.clear();
    off.assign(depth + 1, 0);
    cpn.assign(depth + 1, 1);
    maxc_out.assign(depth + 1, 0);
    std::vector<int> cur = {0, n}, nxt;
    for (int l = 0; l <= depth; ++l) {
        off[l] = (int)bnd.size();
        bnd.insert(bnd.end(), cur.begin(), cur.end());
        int maxc = 0;
        for (size_t j = 0; j + 1 < cur.size(); ++j) maxc = std::max(maxc, cur[j + 1] - cur[j]);
        cpn[l] = std::max(1, (maxc + kBuildChunk - 1) / kBuildChunk);
        maxc_out[l] = maxc;
        if (l == depth) {
            max_leaf = maxc;
            break;
        }
        nxt.clear();
        for (size_t j = 0; j + 1 < cur.size(); ++j) {
            nxt.push_back(cur[j]);
            nxt.push_back(cur[j] + (cur[j + 1] - cur[j]) / 2);
        }
        nxt.push_back(n);
        cur.swap(nxt);
    }
}

// Builds the tree from n x dpad padded points already in c->pts_a.  Without
// need_pts the points are not written in tree order (the tree then has no
// pts: the full-patch path reads only its nodes, order and leaves).
int build_from_padded(annf_ctx *c, int n, int d, int dpad, int depth, bool need_pts = true) {
    std::vector<int> &bnd = c->h_bnd, &off = c->h_off, &cpn = c->h_cpn, &maxc = c->h_maxc;
    // the level bounds depend on (n, depth) alone: computed and uploaded once
    const long long bkey = ((long long)n << 6) | depth;
    const bool same_bounds = c->bnd_key == bkey;
    if (!same_bounds) {
        c->bnd_key = -1;
        level_bounds(n, depth, bnd, off, cpn, maxc, c->bnd_max_leaf);
    }
    const int max_leaf = c->bnd_max_leaf;
    const size_t nodes_max = depth > 0 ? ((size_t)1 << (depth - 1)) : 1;
    size_t chunks_max = 1;
    for (int l = 0; l < depth; ++l) chunks_max = std::max(chunks_max, ((size_t)1 << l) * cpn[l]);
    const size_t nn = depth > 0 ? ((size_t)1 << depth) - 1 : 1;
    RCHK(c->pts_b.ensure((size_t)n * dpad * sizeof(float)));
    RCHK(c->orig_a.ensure((size_t)n * sizeof(int)));
    RCHK(c->orig_b.ensure((size_t)n * sizeof(int)));
    RCHK(c->axis.ensure(nn * sizeof(int)));
    RCHK(c->split.ensure(nn * sizeof(float)));
    RCHK(c->leaves.ensure(((size_t)2 << depth) * sizeof(int)));
    RCHK(c->pli.ensure((size_t)n * sizeof(int)));
    RCHK(c->keys.ensure((size_t)n * sizeof(uint32_t)));
    RCHK(c->eqoff.ensure((size_t)n * sizeof(int)));
    size_t box_words = (size_t)dpad * 2;
    for (int l = 0; l < depth; ++l) box_words = std::max(box_words, level_box_words(1 << l, cpn[l], dpad));
    RCHK(c->box.ensure(box_words * sizeof(uint32_t)));
    RCHK(c->box2.ensure(box_words * sizeof(uint32_t)));
    RCHK(c->nax.ensure(nodes_max * sizeof(int)));
    size_t hist_words = 1;
    for (int l = 0; l < depth; ++l) hist_words = std::max(hist_words, level_hist_words(1 << l, cpn[l]));
    RCHK(c->prefix.ensure(3 * nodes_max * sizeof(uint32_t)));
    RCHK(c->rank.ensure(2 * nodes_max * sizeof(int)));
    RCHK(c->hist.ensure(3 * hist_words * sizeof(uint32_t)));
    RCHK(c->cnt.ensure(chunks_max * sizeof(int2)));
    RCHK(c->cnt_off.ensure(chunks_max * sizeof(int2)));
    RCHK(c->tot.ensure(nodes_max * sizeof(int2)));
    const void *bnd_was = c->bnd.p;
    RCHK(c->bnd.ensure(bnd.size() * sizeof(int)));
    RCHK(c->errflag.ensure(sizeof(int)));
    HIPCHK(hipMemsetAsync(c->errflag.p, 0, sizeof(int), c->stream));
    if (!same_bounds || c->bnd.p != bnd_was) {
        HIPCHK(hipMemcpyAsync(c->bnd.p, bnd.data(), bnd.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        c->bnd_key = bkey;
    }
    BuildScratch s;
    s.pts_a = c->pts_a.as<float>();
    s.pts_b = c->pts_b.as<float>();
    s.orig_a = c->orig_a.as<int>();
    s.orig_b = c->orig_b.as<int>();
    s.keys = c->keys.as<uint32_t>();
    s.eqoff = c->eqoff.as<int>();
    s.box = c->box.as<uint32_t>();
    s.box2 = c->box2.as<uint32_t>();
    s.nax = c->nax.as<int>();
    s.prefix = c->prefix.as<uint32_t>();
    s.rank = c->rank.as<int>();
    s.hist = c->hist.as<uint32_t>();
    s.hist_words = hist_words;
    s.nodes_max = nodes_max;
    s.cnt = c->cnt.as<int2>();
    s.cnt_off = c->cnt_off.as<int2>();
    s.tot = c->tot.as<int2>();
    s.bnd = c->bnd.as<int>();
    s.err = c->errflag.as<int>();
    TreeView t;
    HIPCHK(build_tree(n, d, dpad, depth, c->splitting_type, bnd.data(), off.data(), cpn.data(), maxc.data(), s,
                      c->axis.as<int>(), c->split.as<float>(), c->leaves.as<int>(), c->pli.as<int>(), t,
                      c->stream, need_pts));
    t.max_leaf = max_leaf;
    c->tree = t;
    c->has_tree = true;
    return ANNF_OK;
}

int check_tree_params(annf_ctx *c, int n, int d, int depth) {
    if (!c->inited) return fail(ANNF_EINVAL, "annf_init must be called before fitting");
    if (n < 1) return fail(ANNF_EINVAL, "empty training set");
    if (d < 1 || d > ANNF_MAX_D) return fail(ANNF_EUNSUPPORTED, "dimension %d outside [1, %d]", d, ANNF_MAX_D);
    if (depth < 0 || depth > 30) return fail(ANNF_EINVAL, "tree depth %d outside [0, 30]", depth);
    if ((n >> depth) < 1)
        return fail(ANNF_EINVAL, "tree depth %d leaves empty leaves for %d points", depth, n);
    return ANNF_OK;
}

// Rows pipeline on padded queries in c->qbuf (n_a x dpad).
int run_rows(annf_ctx *c, int n_a, int n_rows, int n_cols, int K, int propagate, int select_best,
             const PatchGeom &ga, const PatchGeom &gb, int *final_dev, float *dist_dev, int *knn_dev) {
    const TreeView &t = c->tree;
    const int nleaves = 1 << t.depth;
    const size_t nk = (size_t)n_a * K;
    RCHK(c->fl_pos.ensure(nk * sizeof(int)));
    RCHK(c->fl_d.ensure((size_t)n_cols * K * sizeof(float)));
    RCHK(c->fl_l.ensure((size_t)n_cols * K * sizeof(int)));
    int *propleaf = nullptr;
    if (propagate) {
        RCHK(c->propleaf.ensure((size_t)t.n * sizeof(int)));
        propleaf = c->propleaf.as<int>();
        HIPCHK(prop_leaves(t, n_a, n_cols, propleaf, c->stream));
    }
    c->mark(3);
    if (n_rows > 1) {
        RCHK(c->natleaf.ensure((size_t)n_a * sizeof(int)));
        RCHK(c->bcount.ensure((size_t)nleaves * sizeof(int)));
        RCHK(c->boffset.ensure(((size_t)nleaves + 1) * sizeof(int)));
        RCHK(c->bcursor.ensure((size_t)nleaves * sizeof(int)));
        RCHK(c->bchunk.ensure(((size_t)nleaves + 1) * sizeof(int)));
        RCHK(c->order.ensure((size_t)n_a * sizeof(int)));
        RCHK(c->sort_keys.ensure((size_t)n_a * sizeof(int)));
        const size_t sort_bytes = sort_scratch_bytes(n_a, nleaves);
        RCHK(c->sort_tmp.ensure(sort_bytes));
        int *natdesc = nullptr;
        if (propagate) {
            // the sweep's leaf descriptor packs a leaf's first tree position
            // as (from << 1) in an int
            if (t.n >= (1 << 30))
                return fail(ANNF_EUNSUPPORTED, "%d tree points: propagation supports fewer than 2^30", t.n);
            RCHK(c->natdesc.ensure((size_t)n_a * sizeof(int)));
            natdesc = c->natdesc.as<int>();
        }
        HIPCHK(natural_leaves(t, c->qbuf.as<float>(), n_a, c->natleaf.as<int>(), nullptr, natdesc, c->stream));
        HIPCHK(bucket_sorted(c->natleaf.as<int>(), n_a, nleaves, leafscan_chunk(), c->sort_keys.as<int>(),
                             c->order.as<int>(), c->boffset.as<int>(), c->bchunk.as<int>(), c->sort_tmp.p,
                             sort_bytes, c->stream));
        if (propagate) {
            RCHK(c->p1d.ensure(nk * sizeof(float)));
            RCHK(c->p1p.ensure(nk * sizeof(int)));
            RCHK(c->p1l.ensure(nk * sizeof(int)));
            HIPCHK(leaf_scan(t, c->qbuf.as<float>(), n_a, c->order.as<int>(), c->boffset.as<int>(),
                             c->bchunk.as<int>(), K, propleaf, c->p1d.as<float>(), c->p1p.as<int>(),
                             c->p1l.as<int>(), c->stream));
        } else {
            // without propagation the natural-leaf list is the row's list
            RCHK(c->p1d.ensure(nk * sizeof(float)));
            HIPCHK(leaf_scan(t, c->qbuf.as<float>(), n_a, c->order.as<int>(), c->boffset.as<int>(),
                             c->bchunk.as<int>(), K, nullptr, c->p1d.as<float>(), c->fl_pos.as<int>(), nullptr,
                             c->stream));
        }
    }
    c->mark(4);
    HIPCHK(exact_search(t, c->qbuf.as<float>(), n_cols, K, 2147483640LL, propleaf, c->fl_d.as<float>(),
                        c->fl_pos.as<int>(), c->fl_l.as<int>(), c->stream));
    c->mark(5);
    if (propagate && n_rows > 1) {
        HIPCHK(row_sweep(t, c->qbuf.as<float>(), n_rows, n_cols, K, propleaf, c->p1d.as<float>(),
                         c->p1p.as<int>(), c->p1l.as<int>(), c->natdesc.as<int>(), c->fl_pos.as<int>(),
                         c->fl_l.as<int>(), c->stream));
    }
    c->mark(6);
    HIPCHK(best_of_k(ga, gb, c->fl_pos.as<int>(), t.orig, n_a, K, select_best, final_dev, dist_dev, knn_dev,
                     c->stream));
    c->mark(7);
    return ANNF_OK;
}

}  // namespace

extern "C" {

const char *annf_last_error(void) { return g_err.c_str(); }

int annf_create(annf_ctx **out, int device) {
    if (!out) return fail(ANNF_EINVAL, "null output handle");
    *out = nullptr;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(ANNF_EINVAL, "device %d not present (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));
    annf_ctx *c = new annf_ctx();
    c->device = device;
    // both streams at the default priority
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
        annf_destroy(c);
        return fail(ANNF_EHIP, "hipStreamCreateWithFlags failed");
    }
    if (hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess) {
        annf_destroy(c);
        return fail(ANNF_EHIP, "hipEventCreate failed");
    }
    for (int i = 0; i <= kStages; ++i) {
        if (hipEventCreate(&c->ev[i]) != hipSuccess) {
            annf_destroy(c);
            return fail(ANNF_EHIP, "hipEventCreate failed");
        }
    }
    if (hipHostMalloc((void **)&c->h_flag, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        c->h_flag = nullptr;
        annf_destroy(c);
        return fail(ANNF_ENOMEM, "hipHostMalloc failed");
    }
    const int rc = scratch_address(device, &c->noalloc_ws);
    if (rc != ANNF_OK) {
        annf_destroy(c);
        return rc;
    }
    *out = c;
    return ANNF_OK;
}

int annf_destroy(annf_ctx *c) {
    if (!c) return ANNF_OK;
    (void)hipSetDevice(c->device);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->release_all();
    for (int i = 0; i <= kStages; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->join_ev) (void)hipEventDestroy(c->join_ev);
    if (c->h_flag) (void)hipHostFree(c->h_flag);
    if (c->blas) (void)rocblas_destroy_handle(c->blas);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ANNF_OK;
}

void *annf_stream(annf_ctx *c) { return c ? (void *)c->stream : nullptr; }

int annf_init(annf_ctx *c, int n_neighbors, int tree_depth, int max_leaves, int num_threads, int splitting_type,
              int verbosity_level) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    // helper.c:30-52
    if (n_neighbors < 1 || n_neighbors > 1000) return fail(ANNF_EINVAL, "k must be > 0 and <= 1000");
    if (tree_depth < 0 || tree_depth > 1000) return fail(ANNF_EINVAL, "tree_depth must positive and <= 1000");
    if (num_threads < 1) return fail(ANNF_EINVAL, "num_threads must be > 0");
    if (verbosity_level < 0 || verbosity_level > 1) return fail(ANNF_EINVAL, "verbosity_level must either 0 or 1");
    // kdtree.c:419-423
    if (splitting_type != ANNF_SPLIT_CYCLIC && splitting_type != ANNF_SPLIT_LONGEST)
        return fail(ANNF_EINVAL, "Unknown splitting type!");
    if (max_leaves < 0) return fail(ANNF_EINVAL, "max_leaves must be a positive integer");
    c->n_neighbors = n_neighbors;
    c->tree_depth = tree_depth;
    c->max_leaves = max_leaves;
    c->num_threads = num_threads;
    c->splitting_type = splitting_type;
    c->verbosity = verbosity_level;
    c->inited = true;
    return ANNF_OK;
}

int annf_fit_dev(annf_ctx *c, const float *X, int n, int d) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    RCHK(set_device(c));
    RCHK(check_tree_params(c, n, d, c->tree_depth));
    const int dpad = dpad_of(d);
    c->has_tree = false;
    RCHK(c->pts_a.ensure((size_t)n * dpad * sizeof(float)));
    HIPCHK(pad_rows(X, n, d, dpad, c->pts_a.as<float>(), c->stream));
    RCHK(build_from_padded(c, n, d, dpad, c->tree_depth));
    return sync_and_check_build(c);
}

int annf_fit(annf_ctx *c, const float *X, int n, int d) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!X) return fail(ANNF_EINVAL, "null Xtrain");
    RCHK(set_device(c));
    RCHK(check_tree_params(c, n, d, c->tree_depth));
    RCHK(c->stage_in.ensure((size_t)n * d * sizeof(float)));
    HIPCHK(hipMemcpyAsync(c->stage_in.p, X, (size_t)n * d * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return annf_fit_dev(c, c->stage_in.as<float>(), n, d);
}

int annf_neighbors(annf_ctx *c, const float *Xtest, int nXtest, int dXtest, float *distances, int ndistances,
                   int ddistances, int *indices, int nindices, int dindices) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!c->has_tree) return fail(ANNF_ENOTREE, "no tree fitted");
    if (!c->tree.pts)
        return fail(ANNF_ENOTREE, "the tree of the last full-patch field holds no points: fit a tree first");
    RCHK(set_device(c));
    const TreeView &t = c->tree;
    const int K = ddistances;
    if (dXtest != t.d) return fail(ANNF_EINVAL, "query dimension %d != tree dimension %d", dXtest, t.d);
    if (ndistances != nXtest || nindices != nXtest || dindices != K)
        return fail(ANNF_EINVAL, "output shapes do not match the queries");
    if (!k_supported(K)) return fail(ANNF_EUNSUPPORTED, "n_neighbors %d outside [1, %d]", K, ANNF_MAX_K);
    if (nXtest == 0) return ANNF_OK;
    const size_t nk = (size_t)nXtest * K;
    RCHK(c->stage_in.ensure((size_t)nXtest * dXtest * sizeof(float)));
    RCHK(c->qbuf.ensure((size_t)nXtest * t.dpad * sizeof(float)));
    RCHK(c->p1d.ensure(nk * sizeof(float)));
    RCHK(c->p1p.ensure(nk * sizeof(int)));
    RCHK(c->final_i.ensure(nk * sizeof(int)));
    RCHK(c->final_d.ensure(nk * sizeof(float)));
    HIPCHK(hipMemcpyAsync(c->stage_in.p, Xtest, (size_t)nXtest * dXtest * sizeof(float), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(pad_rows(c->stage_in.as<float>(), nXtest, dXtest, t.dpad, c->qbuf.as<float>(), c->stream));
    HIPCHK(exact_search(t, c->qbuf.as<float>(), nXtest, K, (long long)c->max_leaves, nullptr, c->p1d.as<float>(),
                        c->p1p.as<int>(), nullptr, c->stream));
    HIPCHK(convert_knn(c->p1p.as<int>(), c->p1d.as<float>(), t.orig, (long long)nk, c->final_i.as<int>(),
                       c->final_d.as<float>(), c->stream));
    HIPCHK(hipMemcpyAsync(indices, c->final_i.p, nk * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(distances, c->final_d.p, nk * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return sync_stream(c);
}

int annf_process_rows(annf_ctx *c, int *finalindices, int nfinalindices, const uint8_t *patchesA, int npatchesA,
                      int dpatchesA, const uint8_t *patchesB, int npatchesB, int dpatchesB, const float *patchesred,
                      int npatchesred, int dpatchesred, int n_rows, int n_cols, int n_neighbors, int propagate,
                      int select_best_nn) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!c->has_tree) return fail(ANNF_ENOTREE, "no tree fitted");
    if (!c->tree.pts)
        return fail(ANNF_ENOTREE, "the tree of the last full-patch field holds no points: fit a tree first");
    RCHK(set_device(c));
    const TreeView &t = c->tree;
    const int K = n_neighbors;
    if (!k_supported(K)) return fail(ANNF_EUNSUPPORTED, "n_neighbors %d outside [1, %d]", K, ANNF_MAX_K);
    if (dpatchesred != t.d) return fail(ANNF_EINVAL, "reduced dimension %d != tree dimension %d", dpatchesred, t.d);
    if (n_rows < 1 || n_cols < 1 || (long long)n_rows * n_cols != npatchesred)
        return fail(ANNF_EINVAL, "n_rows * n_cols (%d * %d) != number of reduced patches %d", n_rows, n_cols,
                    npatchesred);
    if (nfinalindices != npatchesred) return fail(ANNF_EINVAL, "finalindices has %d entries, expected %d",
                                                  nfinalindices, npatchesred);
    if (propagate && npatchesred > t.n)
        return fail(ANNF_EINVAL, "propagation needs at most as many query patches (%d) as tree points (%d)",
                    npatchesred, t.n);
    if (select_best_nn) {
        if (!patchesA || !patchesB) return fail(ANNF_EINVAL, "select_best_nn needs both patch arrays");
        if (npatchesA != npatchesred) return fail(ANNF_EINVAL, "patchesA rows %d != %d", npatchesA, npatchesred);
        if (npatchesB != t.n) return fail(ANNF_EINVAL, "patchesB rows %d != tree points %d", npatchesB, t.n);
        if (dpatchesA != dpatchesB || dpatchesA < 1) return fail(ANNF_EINVAL, "patch dimensions differ");
    }
    const int n_a = npatchesred;
    c->reset_marks();
    c->mark(0);
    RCHK(c->stage_in.ensure((size_t)n_a * t.d * sizeof(float)));
    RCHK(c->qbuf.ensure((size_t)n_a * t.dpad * sizeof(float)));
    RCHK(c->final_i.ensure((size_t)n_a * sizeof(int)));
    HIPCHK(hipMemcpyAsync(c->stage_in.p, patchesred, (size_t)n_a * t.d * sizeof(float), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(pad_rows(c->stage_in.as<float>(), n_a, t.d, t.dpad, c->qbuf.as<float>(), c->stream));
    PatchGeom ga{}, gb{};
    if (select_best_nn) {
        const size_t ba = (size_t)npatchesA * dpatchesA, bb = (size_t)npatchesB * dpatchesB;
        RCHK(c->stage_in2.ensure(ba));
        RCHK(c->stage_in3.ensure(bb));
        HIPCHK(hipMemcpyAsync(c->stage_in2.p, patchesA, ba, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->stage_in3.p, patchesB, bb, hipMemcpyHostToDevice, c->stream));
        ga = PatchGeom{c->stage_in2.as<uint8_t>(), 0, 0, 0, 1, dpatchesA, 0, (long long)ba};
        gb = PatchGeom{c->stage_in3.as<uint8_t>(), 0, 0, 0, 1, dpatchesB, 0, (long long)bb};
    }
    c->mark(1);
    c->mark(2);
    RCHK(run_rows(c, n_a, n_rows, n_cols, K, propagate, select_best_nn, ga, gb, c->final_i.as<int>(), nullptr,
                  nullptr));
    HIPCHK(hipMemcpyAsync(finalindices, c->final_i.p, (size_t)n_a * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RCHK(sync_stream(c));
    c->collect_marks();
    return ANNF_OK;
}

float annf_norm_array(const uint8_t *a, const uint8_t *b, int dim) {
    float s = 0.0f;
    for (int i = 0; i < dim; ++i) {
        float df = (float)a[i] - (float)b[i];
        s += df * df;
    }
    return (float)std::sqrt((double)s);
}

int annf_free_resources(annf_ctx *c) {
    if (!c) return ANNF_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->release_all();
    return ANNF_OK;
}

int annf_tree_info(annf_ctx *c, int *n, int *d, int *depth) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!c->has_tree) return fail(ANNF_ENOTREE, "no tree fitted");
    if (n) *n = c->tree.n;
    if (d) *d = c->tree.d;
    if (depth) *depth = c->tree.depth;
    return ANNF_OK;
}

int annf_tree_export(annf_ctx *c, int *axis, float *split, int *leaves, int *xi_orig, float *xi_pts, int *pli) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!c->has_tree) return fail(ANNF_ENOTREE, "no tree fitted");
    if (xi_pts && !c->tree.pts)
        return fail(ANNF_ENOTREE, "the tree of the last full-patch field holds no points: fit a tree first");
    RCHK(set_device(c));
    const TreeView &t = c->tree;
    const size_t nn = ((size_t)1 << t.depth) - 1;
    if (axis && nn) HIPCHK(hipMemcpyAsync(axis, t.axis, nn * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (split && nn) HIPCHK(hipMemcpyAsync(split, t.split, nn * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (leaves)
        HIPCHK(hipMemcpyAsync(leaves, t.leaves, ((size_t)2 << t.depth) * sizeof(int), hipMemcpyDeviceToHost,
                              c->stream));
    if (xi_orig) HIPCHK(hipMemcpyAsync(xi_orig, t.orig, (size_t)t.n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (pli) HIPCHK(hipMemcpyAsync(pli, t.pli, (size_t)t.n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (xi_pts)
        HIPCHK(hipMemcpy2DAsync(xi_pts, (size_t)t.d * sizeof(float), t.pts, (size_t)t.dpad * sizeof(float),
                                (size_t)t.d * sizeof(float), (size_t)t.n, hipMemcpyDeviceToHost, c->stream));
    return sync_stream(c);
}

int annf_tree_import(annf_ctx *c, int n, int d, int depth, const int *axis, const float *split, const int *leaves,
                     const int *xi_orig, const float *xi_pts, const int *pli) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!axis || !split || !leaves || !xi_orig || !xi_pts || !pli) return fail(ANNF_EINVAL, "null tree array");
    RCHK(set_device(c));
    RCHK(check_tree_params(c, n, d, depth));
    if (depth != c->tree_depth)
        return fail(ANNF_EINVAL, "tree depth %d differs from the handle's depth %d", depth, c->tree_depth);
    c->has_tree = false;
    // the leaf ranges are fixed by n and depth (the left child of c points gets c/2);
    // local bounds: the handle's cached build bounds stay as they are
    std::vector<int> bnd, off, cpn, maxc;
    int max_leaf = 0;
    level_bounds(n, depth, bnd, off, cpn, maxc, max_leaf);
    const int nleaf = 1 << depth;
    for (int l = 0; l < nleaf; ++l)
        if (leaves[2 * l] != bnd[off[depth] + l] || leaves[2 * l + 1] != bnd[off[depth] + l + 1])
            return fail(ANNF_EINVAL, "leaf %d range [%d, %d) is not the median split of %d points", l,
                        leaves[2 * l], leaves[2 * l + 1], n);
    const size_t nn = ((size_t)1 << depth) - 1;
    for (size_t i = 0; i < nn; ++i)
        if (axis[i] < 0 || axis[i] >= d) return fail(ANNF_EINVAL, "node %zu splits on axis %d", i, axis[i]);
    std::vector<char> seen(n, 0);
    for (int l = 0; l < nleaf; ++l)
        for (int i = leaves[2 * l]; i < leaves[2 * l + 1]; ++i) {
            const int o = xi_orig[i];
            if (o < 0 || o >= n || seen[o]) return fail(ANNF_EINVAL, "xi_orig is not a permutation of [0, %d)", n);
            seen[o] = 1;
            if (pli[o] != l) return fail(ANNF_EINVAL, "pli[%d] = %d, the point lies in leaf %d", o, pli[o], l);
        }
    const int dpad = dpad_of(d);
    RCHK(c->pts_a.ensure((size_t)n * dpad * sizeof(float)));
    RCHK(c->orig_a.ensure((size_t)n * sizeof(int)));
    RCHK(c->axis.ensure(std::max<size_t>(nn, 1) * sizeof(int)));
    RCHK(c->split.ensure(std::max<size_t>(nn, 1) * sizeof(float)));
    RCHK(c->leaves.ensure((size_t)2 * nleaf * sizeof(int)));
    RCHK(c->pli.ensure((size_t)n * sizeof(int)));
    HIPCHK(hipMemsetAsync(c->pts_a.p, 0, (size_t)n * dpad * sizeof(float), c->stream));
    HIPCHK(hipMemcpy2DAsync(c->pts_a.p, (size_t)dpad * sizeof(float), xi_pts, (size_t)d * sizeof(float),
                            (size_t)d * sizeof(float), (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->orig_a.p, xi_orig, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (nn) {
        HIPCHK(hipMemcpyAsync(c->axis.p, axis, nn * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->split.p, split, nn * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipMemcpyAsync(c->leaves.p, leaves, (size_t)2 * nleaf * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->pli.p, pli, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    RCHK(sync_stream(c));
    TreeView t;
    t.n = n;
    t.d = d;
    t.dpad = dpad;
    t.depth = depth;
    t.max_leaf = max_leaf;
    t.pts = c->pts_a.as<float>();
    t.orig = c->orig_a.as<int>();
    t.axis = c->axis.as<int>();
    t.split = c->split.as<float>();
    t.leaves = c->leaves.as<int>();
    t.pli = c->pli.as<int>();
    c->tree = t;
    c->has_tree = true;
    return ANNF_OK;
}

}  // extern "C"

namespace {

// basis: d x D, float32 (basis_f64 == 0) or float64; the FP32 table holds it
// rounded once, the FP64 table as given
__global__ void k_prep_basis(const void *basis, int basis_f64, const double *mean, int d, int D, int dpad,
                             float *bt32, float *m32, double *bt64, double *m64) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < D * dpad) {
        int f = i / dpad, k = i - f * dpad;
        double v = 0.0;
        if (k < d)
            v = basis_f64 ? static_cast<const double *>(basis)[(size_t)k * D + f]
                          : (double)static_cast<const float *>(basis)[(size_t)k * D + f];
        bt32[i] = (float)v;
        bt64[i] = v;
    }
    if (i < D) {
        m32[i] = (float)mean[i];
        m64[i] = mean[i];
    }
}

__global__ void k_unpad(const float *src, long long n, int d, int dpad, float *dst) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * d) return;
    long long r = i / d;
    int c = (int)(i - r * d);
    dst[i] = src[r * dpad + c];
}

// Validates a field request, projects A (into c->qbuf) and B, and builds the
// tree over B.  Stage marks 0..3.
struct FieldShape {
    int p = 0, nr_a = 0, nc_a = 0;
    long long n_a = 0, n_b = 0;
    PatchGeom ga{}, gb{};
};

int project_and_build(annf_ctx *c, const uint8_t *img_a, int ha, int wa, const uint8_t *img_b, int hb, int wb,
                      int psize, const void *basis, const double *mean, int d, int projection, int tree_depth,
                      int K, float *red_a_out, float *red_b_out, FieldShape &fs,
                      const std::function<int(hipStream_t)> &side_extra = nullptr, bool need_pts = true) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!img_a || !img_b) return fail(ANNF_EINVAL, "null image pointer");
    RCHK(set_device(c));
    const int p = psize;
    if (p < 1 || ha < p || wa < p || hb < p || wb < p) return fail(ANNF_EINVAL, "patch size %d larger than an image", p);
    if (d < 1 || d > ANNF_MAX_D) return fail(ANNF_EUNSUPPORTED, "dimension %d outside [1, %d]", d, ANNF_MAX_D);
    if (projection != ANNF_PROJ_PCA32 && projection != ANNF_PROJ_PCA64 && projection != ANNF_PROJ_ENCODER)
        return fail(ANNF_EINVAL, "projection %d is not ANNF_PROJ_PCA32, _PCA64 or _ENCODER", projection);
    const bool enc = projection == ANNF_PROJ_ENCODER;
    if (!enc && (!basis || !mean)) return fail(ANNF_EINVAL, "null basis or mean");
    if (enc && (!c->has_encoder || c->enc_p != p || c->enc_d != d))
        return fail(ANNF_EINVAL, "no encoder for patch size %d and dimension %d on this handle (annf_set_encoder)",
                    p, d);
    if (!k_supported(K)) return fail(ANNF_EUNSUPPORTED, "n_neighbors %d outside [1, %d]", K, ANNF_MAX_K);
    const int D = p * p * 3;
    const int nr_a = ha - p + 1, nc_a = wa - p + 1, nr_b = hb - p + 1, nc_b = wb - p + 1;
    const long long n_a = (long long)nr_a * nc_a, n_b = (long long)nr_b * nc_b;
    if (n_a > 0x7FFFFFFFLL || n_b > 0x7FFFFFFFLL) return fail(ANNF_EINVAL, "too many patches");
    if (!c->inited) RCHK(annf_init(c, K, tree_depth, 2147483640, 1, ANNF_SPLIT_LONGEST, 0));
    c->tree_depth = tree_depth;
    RCHK(check_tree_params(c, (int)n_b, d, tree_depth));
    const int dpad = dpad_of(d);
    const PatchGeom ga{img_a, 1, nc_a, wa, p, 3 * p, 3LL * wa, 3LL * wa * ha};
    const PatchGeom gb{img_b, 1, nc_b, wb, p, 3 * p, 3LL * wb, 3LL * wb * hb};
    c->reset_marks();
    c->mark(0);
    if (!enc) {
        RCHK(c->basis32.ensure((size_t)D * dpad * sizeof(float)));
        RCHK(c->mean32.ensure((size_t)D * sizeof(float)));
        RCHK(c->basis64.ensure((size_t)D * dpad * sizeof(double)));
        RCHK(c->mean64.ensure((size_t)D * sizeof(double)));
        hipLaunchKernelGGL(k_prep_basis, dim3((D * dpad + 255) / 256), dim3(256), 0, c->stream, basis,
                           (int)(projection == ANNF_PROJ_PCA64), mean, d, D, dpad, c->basis32.as<float>(), c->mean32.as<float>(), c->basis64.as<double>(),
                           c->mean64.as<double>());
        HIPCHK(hipGetLastError());
        if (projection == ANNF_PROJ_PCA32) {
            RCHK(c->proj_pack.ensure(projection_pack_floats(p, dpad) * sizeof(float)));
            HIPCHK(pack_projection(c->basis32.as<float>(), c->mean32.as<float>(), p, dpad, c->proj_pack.as<float>(),
                                   c->stream));
        }
    }
    auto project = [&](const uint8_t *img, int h, int w, const PatchGeom &g, long long n, float *out,
                       hipStream_t st, int per_cu) -> int {
        if (enc) {
            HIPCHK(encode_patches(g, n, c->enc_w1.as<__bf16>(), c->enc_b1.as<float>(), c->enc_w2.as<__bf16>(),
                                  c->enc_b2.as<float>(), c->enc_w3.as<__bf16>(), c->enc_b3.as<float>(), d, dpad, out,
                                  st, per_cu));
        } else {
            HIPCHK(project_patches(img, h, w, p, c->basis32.as<float>(), c->mean32.as<float>(),
                                   c->basis64.as<double>(), c->mean64.as<double>(), d, dpad,
                                   projection == ANNF_PROJ_PCA64, out, c->proj_pack.as<float>(), per_cu, st));
        }
        return ANNF_OK;
    };
    RCHK(c->qbuf.ensure((size_t)n_a * dpad * sizeof(float)));
    c->has_tree = false;
    RCHK(c->pts_a.ensure((size_t)n_b * dpad * sizeof(float)));
    // B is projected on `stream`; then A (the queries) is projected on the
    // side stream while `stream` builds B's tree, and `stream` waits for A
    // before anything reads it
    RCHK(project(img_b, hb, wb, gb, n_b, c->pts_a.as<float>(), c->stream, 2));
    if (red_b_out) {
        hipLaunchKernelGGL(k_unpad, dim3((unsigned)((n_b * d + 255) / 256)), dim3(256), 0, c->stream,
                           c->pts_a.as<float>(), n_b, d, dpad, red_b_out);
        HIPCHK(hipGetLastError());
    }
    c->mark(1);
    HIPCHK(hipEventRecord(c->fork_ev, c->stream));
    HIPCHK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    // one encoder block per CU: the build's short kernels keep room beside it
    int rc = project(img_a, ha, wa, ga, n_a, c->qbuf.as<float>(), c->side, 1);
    // work that needs only the images (queued behind A's projection, beside the build)
    if (rc == ANNF_OK && side_extra) rc = side_extra(c->side);
    if (rc == ANNF_OK && hipEventRecord(c->join_ev, c->side) != hipSuccess) rc = fail(ANNF_EHIP, "hipEventRecord failed");
    if (rc != ANNF_OK) {
        (void)hipStreamSynchronize(c->side);
        return rc;
    }
    c->mark(2);
    rc = build_from_padded(c, (int)n_b, d, dpad, tree_depth, need_pts);
    if (hipStreamWaitEvent(c->stream, c->join_ev, 0) != hipSuccess && rc == ANNF_OK)
        rc = fail(ANNF_EHIP, "hipStreamWaitEvent failed");
    if (rc != ANNF_OK) {
        (void)hipStreamSynchronize(c->side);
        return rc;
    }
    if (red_a_out) {
        hipLaunchKernelGGL(k_unpad, dim3((unsigned)((n_a * d + 255) / 256)), dim3(256), 0, c->stream,
                           c->qbuf.as<float>(), n_a, d, dpad, red_a_out);
        HIPCHK(hipGetLastError());
    }
    fs.p = p;
    fs.nr_a = nr_a;
    fs.nc_a = nc_a;
    fs.n_a = n_a;
    fs.n_b = n_b;
    fs.ga = ga;
    fs.gb = gb;
    return ANNF_OK;
}

}  // namespace

extern "C" {

int annf_field_dev(annf_ctx *c, const uint8_t *img_a, int ha, int wa, const uint8_t *img_b, int hb, int wb,
                   int psize, const void *basis, const double *mean, int d, int projection, int tree_depth,
                   int n_neighbors, int propagate, int select_best_nn, int *nn_out, float *dist_out, int *knn_out,
                   float *red_a_out, float *red_b_out) {
    if (!nn_out) return fail(ANNF_EINVAL, "null required pointer");
    if (propagate && (long long)(ha - psize + 1) * (wa - psize + 1) > (long long)(hb - psize + 1) * (wb - psize + 1))
        return fail(ANNF_EINVAL, "propagation needs at most as many A patches as B patches");
    FieldShape fs;
    RCHK(project_and_build(c, img_a, ha, wa, img_b, hb, wb, psize, basis, mean, d, projection, tree_depth,
                           n_neighbors, red_a_out, red_b_out, fs));
    RCHK(run_rows(c, (int)fs.n_a, fs.nr_a, fs.nc_a, n_neighbors, propagate, select_best_nn, fs.ga, fs.gb, nn_out,
                  dist_out, knn_out));
    RCHK(sync_and_check_build(c));
    c->collect_marks();
    return ANNF_OK;
}

int annf_field_mm_dev(annf_ctx *c, const uint8_t *img_a, int ha, int wa, const uint8_t *img_b, int hb, int wb,
                      int psize, const void *basis, const double *mean, int d, int projection, int tree_depth,
                      int n_neighbors, int nprobe, int precision, int shortlist, int propagate, int *nn_out,
                      float *dist_out, int *knn_out, float *red_a_out, float *red_b_out) {
    if (!nn_out) return fail(ANNF_EINVAL, "null required pointer");
    if (!leafbf16_supported(psize) || !leaffp8_supported(psize))
        return fail(ANNF_EUNSUPPORTED, "patch size %d: the full-patch leaf passes take 1..12", psize);
    if (precision != ANNF_PREC_BF16 && precision != ANNF_PREC_FP8)
        return fail(ANNF_EINVAL, "precision %d is not ANNF_PREC_BF16 or ANNF_PREC_FP8", precision);
    const int K = n_neighbors;
    if (K < 1 || K > 32) return fail(ANNF_EUNSUPPORTED, "n_neighbors %d: the full-patch variants take 1..32", K);
    if (precision == ANNF_PREC_FP8 && (shortlist < K || shortlist > 32))
        return fail(ANNF_EUNSUPPORTED, "shortlist %d outside [n_neighbors, 32]", shortlist);
    if (tree_depth < 0 || tree_depth > 30 || nprobe < 1 || nprobe > (1 << tree_depth) || (nprobe & (nprobe - 1)))
        return fail(ANNF_EINVAL, "nprobe %d is not a power of two in [1, 2^tree_depth]", nprobe);
    if (precision == ANNF_PREC_FP8 && nprobe > 32)
        return fail(ANNF_EUNSUPPORTED, "nprobe %d: the FP8 shortlist merges at most 32 probed leaves", nprobe);
    if (propagate != 0 && wa != wb)
        return fail(ANNF_EINVAL, "propagation needs images of equal width (base.c:271)");
    // propagate < 0: the reference's row chain (rows in order, each from the
    // final lists of the row above), exact BF16 only
    const bool chain = propagate < 0;
    if (chain && precision != ANNF_PREC_BF16)
        return fail(ANNF_EUNSUPPORTED, "the row chain (propagate < 0) runs with ANNF_PREC_BF16");
    if (propagate < -1000000) return fail(ANNF_EINVAL, "propagate %d out of range", propagate);
    if (precision == ANNF_PREC_BF16 && (2 * cimg_elems(ha, wa) >= (1LL << 31) || 2 * cimg_elems(hb, wb) >= (1LL << 31)))
        return fail(ANNF_EUNSUPPORTED, "the BF16 leaf k-NN needs images of fewer than 2^30 elements");
    const int fp8 = precision == ANNF_PREC_FP8;
    // the exact BF16 pass reads both images as centred BF16 with per-patch
    // norms: made on the side stream while the tree is built
    auto prepare_bf16 = [&](hipStream_t st) -> int {
        const long long na_ = (long long)(ha - psize + 1) * (wa - psize + 1);
        const long long nb_ = (long long)(hb - psize + 1) * (wb - psize + 1);
        RCHK(c->cimg_a.ensure(2 * (size_t)cimg_elems(ha, wa) * sizeof(uint16_t)));
        RCHK(c->cimg_b.ensure(2 * (size_t)cimg_elems(hb, wb) * sizeof(uint16_t)));
        RCHK(c->pnorm_a.ensure((size_t)na_ * sizeof(uint32_t)));
        RCHK(c->pnorm_b.ensure((size_t)nb_ * sizeof(uint32_t)));
        RCHK(c->pnorm_tmp.ensure((size_t)std::max((long long)ha * (wa - psize + 1), (long long)hb * (wb - psize + 1)) *
                                 sizeof(uint32_t)));
        HIPCHK(center_image(img_a, ha, wa, c->cimg_a.as<uint16_t>(), st));
        HIPCHK(center_image(img_b, hb, wb, c->cimg_b.as<uint16_t>(), st));
        HIPCHK(patch_norms_centered(img_a, ha, wa, psize, c->pnorm_tmp.as<uint32_t>(), c->pnorm_a.as<uint32_t>(), st));
        HIPCHK(patch_norms_centered(img_b, hb, wb, psize, c->pnorm_tmp.as<uint32_t>(), c->pnorm_b.as<uint32_t>(), st));
        return ANNF_OK;
    };
    // the FP8 pass reads both images as per-patch E4M3 codes: made on the
    // side stream while the tree is built
    auto prepare_fp8 = [&](hipStream_t st) -> int {
        const long long na_ = (long long)(ha - psize + 1) * (wa - psize + 1);
        const long long nb_ = (long long)(hb - psize + 1) * (wb - psize + 1);
        const size_t dq = (size_t)leaffp8_dq(psize);
        RCHK(c->f8_qa.ensure((size_t)na_ * dq));
        RCHK(c->f8_qb.ensure((size_t)nb_ * dq));
        RCHK(c->f8_na.ensure((size_t)na_ * sizeof(float)));
        RCHK(c->f8_ia.ensure((size_t)na_ * sizeof(float)));
        RCHK(c->f8_nb.ensure((size_t)nb_ * sizeof(float)));
        RCHK(c->f8_ib.ensure((size_t)nb_ * sizeof(float)));
        HIPCHK(patch_fp8(img_a, ha, wa, psize, c->f8_qa.as<uint8_t>(), c->f8_na.as<float>(), c->f8_ia.as<float>(),
                         st));
        HIPCHK(patch_fp8(img_b, hb, wb, psize, c->f8_qb.as<uint8_t>(), c->f8_nb.as<float>(), c->f8_ib.as<float>(),
                         st));
        // centred patch norms for the exact re-rank
        RCHK(c->pnorm_a.ensure((size_t)na_ * sizeof(uint32_t)));
        RCHK(c->pnorm_b.ensure((size_t)nb_ * sizeof(uint32_t)));
        RCHK(c->pnorm_tmp.ensure((size_t)std::max((long long)ha * (wa - psize + 1), (long long)hb * (wb - psize + 1)) *
                                 sizeof(uint32_t)));
        HIPCHK(patch_norms_centered(img_a, ha, wa, psize, c->pnorm_tmp.as<uint32_t>(), c->pnorm_a.as<uint32_t>(), st));
        HIPCHK(patch_norms_centered(img_b, hb, wb, psize, c->pnorm_tmp.as<uint32_t>(), c->pnorm_b.as<uint32_t>(), st));
        return ANNF_OK;
    };
    FieldShape fs;
    RCHK(project_and_build(c, img_a, ha, wa, img_b, hb, wb, psize, basis, mean, d, projection, tree_depth, K,
                           red_a_out, red_b_out, fs,
                           fp8 ? std::function<int(hipStream_t)>(prepare_fp8)
                               : std::function<int(hipStream_t)>(prepare_bf16),
                           false));
    const TreeView &t = c->tree;
    const int nleaves = 1 << t.depth;
    const int n_a = (int)fs.n_a;
    const bool prop = propagate != 0 && fs.nr_a > 1;
    // the first prop_k entries of the list above propagate (K: all, base.c:255-282)
    const int prop_k = std::min(chain ? -propagate : propagate, K);
    const int nc_a = fs.nc_a, nr_a = fs.nr_a;
    const size_t n_pairs = prop ? (size_t)(chain ? nc_a : n_a) * prop_k : 0;
    c->mark(3);
    RCHK(c->natleaf.ensure((size_t)n_a * sizeof(int)));
    RCHK(c->bcount.ensure((size_t)nleaves * sizeof(int)));
    RCHK(c->boffset.ensure(((size_t)nleaves + 1) * sizeof(int)));
    RCHK(c->bcursor.ensure((size_t)nleaves * sizeof(int)));
    RCHK(c->bchunk.ensure(((size_t)nleaves + 1) * sizeof(int)));
    RCHK(c->order.ensure(std::max((size_t)n_a, n_pairs) * sizeof(int)));
    RCHK(c->mm_keys1.ensure((size_t)n_a * K * sizeof(uint64_t)));
    const size_t n_sort = std::max((size_t)n_a, n_pairs);
    RCHK(c->sort_keys.ensure(n_sort * sizeof(int)));
    const size_t sort_bytes = sort_scratch_bytes((int)n_sort, nleaves);
    RCHK(c->sort_tmp.ensure(sort_bytes));
    int *skeys = c->sort_keys.as<int>();
    RCHK(c->mm_blkleaf.ensure(leafbf16_scratch_bytes((int)n_sort, nleaves)));
    void *blk_info = c->mm_blkleaf.p;
    const int chunk = leafbf16_chunk();
    RCHK(c->mm_ptorder.ensure((size_t)fs.n_b * sizeof(int)));
    HIPCHK(leaf_point_order(t, c->mm_ptorder.as<int>(), c->stream));
    const int *pts_order = c->mm_ptorder.as<int>();
    // FP8: position bits of the approximate keys (the largest leaf) and the
    // per-(slot, leaf) lists
    int pb = 1;
    while (pb < 24 && (1 << pb) < c->bnd_max_leaf) ++pb;
    if (fp8 && c->bnd_max_leaf > (1 << pb))
        return fail(ANNF_EUNSUPPORTED, "a leaf of %d points: the FP8 keys hold positions below 2^24", c->bnd_max_leaf);
    const int sp8 = leaffp8_stride(shortlist);
    if (fp8) {
        RCHK(c->f8_keys.ensure(std::max((size_t)n_a * nprobe, n_pairs) * sp8 * sizeof(uint64_t)));
    }
    // FP8 leaf pass of nq bucketed slots (slot s: A patch a0 + s / sdiv)
    auto fp8_pass = [&](int nq, int np_, int sdiv) -> int {
        HIPCHK(leaf_fp8(t, c->f8_qa.as<uint8_t>(), c->f8_na.as<float>(), c->f8_ia.as<float>(), c->f8_qb.as<uint8_t>(),
                        c->f8_nb.as<float>(), c->f8_ib.as<float>(), psize, pts_order, c->order.as<int>(),
                        c->boffset.as<int>(), c->bchunk.as<int>(), blk_info, nq, np_, shortlist, sdiv, 0, pb,
                        c->f8_keys.as<uint64_t>(), c->stream));
        return ANNF_OK;
    };
    auto leaf_pass = [&](int nq, int np_, int sdiv, uint64_t *keys, int *knn, int *nn, float *dist) -> int {
        if (fp8) {
            // approximate lists, each query's shortlist, its exact re-rank
            RCHK(fp8_pass(nq, np_, sdiv));
            // queries in bucket order: a leaf's queries share its points
            HIPCHK(fp8_select(c->f8_keys.as<uint64_t>(), shortlist, 0, n_a, c->order.as<int>(), np_, nullptr,
                              fs.ga, fs.gb, c->pnorm_a.as<uint32_t>(), c->pnorm_b.as<uint32_t>(), nullptr, K, keys,
                              knn, nn, dist, c->stream));
        } else {
            HIPCHK(leaf_bf16(t, c->cimg_a.as<uint16_t>(), wa, c->cimg_b.as<uint16_t>(), wb, psize,
                             c->pnorm_a.as<uint32_t>(), c->pnorm_b.as<uint32_t>(), pts_order, c->order.as<int>(),
                             c->boffset.as<int>(), c->bchunk.as<int>(), blk_info, nq, np_, K, sdiv, keys, knn, nn,
                             dist, c->stream, ha, hb));
        }
        return ANNF_OK;
    };
    HIPCHK(natural_leaves(t, c->qbuf.as<float>(), n_a, c->natleaf.as<int>(), nullptr, nullptr, c->stream));
    HIPCHK(bucket_sorted(c->natleaf.as<int>(), n_a, nleaves, chunk, skeys, c->order.as<int>(),
                         c->boffset.as<int>(), c->bchunk.as<int>(), c->sort_tmp.p, sort_bytes, c->stream));
    c->mark(4);
    RCHK(leaf_pass(n_a, nprobe, 1, c->mm_keys1.as<uint64_t>(), prop ? nullptr : knn_out, prop ? nullptr : nn_out,
                   prop ? nullptr : dist_out));
    c->mark(5);
    uint64_t *keys2 = nullptr;
    if (prop) {
        // the propagated leaves, bucketed like the queries, scanned by the same kernel
        RCHK(c->mm_pairleaf.ensure(n_pairs * sizeof(int)));
        if (!fp8) {
            RCHK(c->mm_keys2.ensure(n_pairs * K * sizeof(uint64_t)));
            keys2 = c->mm_keys2.as<uint64_t>();
        }
        // pairs of queries q0 .. q0 + nq - 1 from `above` (keys by query),
        // scanned into keys2
        auto pair_pass = [&](const uint64_t *above, int q0, int nq) -> int {
            const size_t np = (size_t)nq * prop_k;
            HIPCHK(prop_pairs(above, q0, nq, K, prop_k, nprobe, c->natleaf.as<int>(), t.pli, nc_a, (int)fs.n_b,
                              c->mm_pairleaf.as<int>(), nullptr, c->stream));
            HIPCHK(bucket_sorted(c->mm_pairleaf.as<int>(), (int)np, nleaves, chunk, skeys, c->order.as<int>(),
                                 c->boffset.as<int>(), c->bchunk.as<int>(), c->sort_tmp.p, sort_bytes, c->stream));
            if (fp8) {
                // the pairs' approximate lists, then one shortlist and exact
                // re-rank per query, merged with its first-pass keys
                RCHK(fp8_pass((int)np, 1, prop_k));
                HIPCHK(fp8_select(c->f8_keys.as<uint64_t>(), shortlist, q0, nq, nullptr, prop_k,
                                  c->mm_pairleaf.as<int>(), fs.ga, fs.gb, c->pnorm_a.as<uint32_t>(),
                                  c->pnorm_b.as<uint32_t>(), c->mm_keys1.as<uint64_t>(), K, nullptr, knn_out,
                                  nn_out, dist_out, c->stream));
                return ANNF_OK;
            }
            HIPCHK(leaf_bf16(t, c->cimg_a.as<uint16_t>(), wa, c->cimg_b.as<uint16_t>(), wb, psize,
                             c->pnorm_a.as<uint32_t>(), c->pnorm_b.as<uint32_t>(), pts_order, c->order.as<int>(),
                             c->boffset.as<int>(), c->bchunk.as<int>(), blk_info, (int)np, 1, K, prop_k, keys2,
                             nullptr, nullptr, nullptr, c->stream, ha, hb, q0));
            return ANNF_OK;
        };
        if (!chain && fp8) {
            RCHK(pair_pass(c->mm_keys1.as<uint64_t>(), 0, n_a));
        } else if (!chain) {
            RCHK(pair_pass(c->mm_keys1.as<uint64_t>(), 0, n_a));
            HIPCHK(finish_field(c->mm_keys1.as<uint64_t>(), keys2, c->mm_pairleaf.as<int>(), 0, n_a, K, prop_k,
                                knn_out, nn_out, dist_out, nullptr, c->stream));
        } else {
            // row by row: row y's pairs come from row y - 1's final keys
            // (row 0's are its first-pass keys), the reference's dependence
            RCHK(c->mm_fkeys.ensure((size_t)n_a * K * sizeof(uint64_t)));
            uint64_t *fk = c->mm_fkeys.as<uint64_t>();
            HIPCHK(finish_field(c->mm_keys1.as<uint64_t>(), nullptr, nullptr, 0, nc_a, K, 1, knn_out, nn_out,
                                dist_out, fk, c->stream));
            for (int y = 1; y < nr_a; ++y) {
                RCHK(pair_pass(fk, y * nc_a, nc_a));
                HIPCHK(finish_field(c->mm_keys1.as<uint64_t>(), keys2, c->mm_pairleaf.as<int>(), y * nc_a, nc_a, K,
                                    prop_k, knn_out, nn_out, dist_out, fk, c->stream));
            }
        }
    }
    c->mark(6);
    RCHK(sync_and_check_build(c));
    c->collect_marks();
    return ANNF_OK;
}

int annf_exact_field_dev(annf_ctx *c, const uint8_t *img_a, int ha, int wa, const uint8_t *img_b, int hb, int wb,
                         int psize, const int *query_idx, int nq, int k, int *out_idx, int *out_d2) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!img_a || !img_b || (nq > 0 && (!query_idx || !out_idx || !out_d2))) return fail(ANNF_EINVAL, "null pointer");
    RCHK(set_device(c));
    const int p = psize;
    if (p < 1 || ha < p || wa < p || hb < p || wb < p) return fail(ANNF_EINVAL, "patch size %d larger than an image", p);
    if (k < 1 || k > 32) return fail(ANNF_EUNSUPPORTED, "k %d outside [1, 32]", k);
    const int D = p * p * 3, dq = exact_dq(p);
    if (dq == 0) return fail(ANNF_EUNSUPPORTED, "patch of %d bytes: at most 512 supported", D);
    const long long n_a = (long long)(ha - p + 1) * (wa - p + 1), n_b = (long long)(hb - p + 1) * (wb - p + 1);
    if (n_b > 0x7FFFFFFFLL || n_a > 0x7FFFFFFFLL) return fail(ANNF_EINVAL, "too many patches");
    if ((long long)k > n_b) return fail(ANNF_EINVAL, "k %d exceeds the %lld B patches", k, n_b);
    // the sampled A patches' rows, every B patch's row, their norms
    const size_t nqa = (size_t)(nq > 0 ? nq : 1);
    RCHK(c->ev_am.ensure(nqa * dq));
    RCHK(c->ev_an.ensure(nqa * sizeof(uint32_t)));
    RCHK(c->ev_bm.ensure((size_t)n_b * dq));
    RCHK(c->ev_bn.ensure((size_t)n_b * sizeof(uint32_t)));
    const int splits = exact_splits(n_b);
    RCHK(c->ev_part.ensure(nqa * splits * k * sizeof(uint64_t)));
    HIPCHK(patch_matrix(img_a, ha, wa, p, query_idx, nq, c->ev_am.as<uint8_t>(), dq, c->stream));
    HIPCHK(patch_norms(c->ev_am.as<uint8_t>(), nq, dq, c->ev_an.as<uint32_t>(), c->stream));
    HIPCHK(patch_matrix(img_b, hb, wb, p, nullptr, n_b, c->ev_bm.as<uint8_t>(), dq, c->stream));
    // every B patch's centred norm from box sums over the image
    RCHK(c->pnorm_tmp.ensure((size_t)hb * (wb - p + 1) * sizeof(uint32_t)));
    HIPCHK(patch_norms_centered(img_b, hb, wb, p, c->pnorm_tmp.as<uint32_t>(), c->ev_bn.as<uint32_t>(), c->stream));
    HIPCHK(exact_sample(c->ev_am.as<uint8_t>(), c->ev_an.as<uint32_t>(), nq, c->ev_bm.as<uint8_t>(),
                        c->ev_bn.as<uint32_t>(), n_b, p, dq, k, c->ev_part.as<uint64_t>(), out_idx, out_d2,
                        c->stream));
    return sync_stream(c);
}

int annf_mlp_forward_dev(annf_ctx *c, const uint8_t *x, int n, int D, const float *params, int h1, int h2, int d,
                         float *out) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (n < 0 || D < 1 || h1 < 1 || h2 < 1 || d < 1) return fail(ANNF_EINVAL, "bad MLP shape");
    if (n > 0 && (!x || !params || !out)) return fail(ANNF_EINVAL, "null pointer");
    RCHK(set_device(c));
    if (n == 0) return ANNF_OK;
    RCHK(c->mlp_ws.ensure(mlp_workspace_floats(n, h1, h2, d) * sizeof(float)));
    HIPCHK(mlp_forward(x, n, D, params, h1, h2, d, c->mlp_ws.as<float>(), out, c->stream));
    return sync_stream(c);
}

int annf_mlp_triplet_grad_dev(annf_ctx *c, const uint8_t *x, int B, int D, const float *params, int h1, int h2,
                              int d, float margin, float *grads, float *loss) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (B < 1 || D < 1 || h1 < 1 || h2 < 1 || d < 1) return fail(ANNF_EINVAL, "bad MLP shape or empty batch");
    if (!x || !params || !grads) return fail(ANNF_EINVAL, "null pointer");
    if (3LL * B > 0x7FFFFFFFLL) return fail(ANNF_EINVAL, "batch too large");
    RCHK(set_device(c));
    RCHK(c->mlp_ws.ensure(mlp_workspace_floats(3LL * B, h1, h2, d) * sizeof(float)));
    HIPCHK(mlp_triplet_grad(x, B, D, params, h1, h2, d, margin, grads, loss, c->mlp_ws.as<float>(), c->stream));
    return sync_stream(c);
}

int annf_mlp_triplet_grad_noalloc_dev(annf_ctx *c, const uint8_t *x, int B, int D, const float *params, int h1,
                                      int h2, int d, float margin, float *grads, float *loss) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (B < 1 || D < 1 || h1 < 1 || h2 < 1 || d < 1) return fail(ANNF_EINVAL, "bad MLP shape or empty batch");
    if (!x || !params || !grads) return fail(ANNF_EINVAL, "null pointer");
    if (3LL * B > 0x7FFFFFFFLL) return fail(ANNF_EINVAL, "batch too large");
    const size_t need = mlp_workspace_floats(3LL * B, h1, h2, d);
    if (need > kNoAllocFloats)
        return fail(ANNF_EUNSUPPORTED, "batch %d of a %d-%d-%d network needs %zu workspace floats; the "
                    "allocation-free entry holds %zu (use annf_mlp_triplet_grad_dev)", B, h1, h2, d, need,
                    kNoAllocFloats);
    if (!c->noalloc_ws) return fail(ANNF_EHIP, "no workspace address for device %d", c->device);
    RCHK(set_device(c));
    // every workspace word the kernels read is written earlier in the same
    // call, so nothing is cleared
    HIPCHK(mlp_triplet_grad(x, B, D, params, h1, h2, d, margin, grads, loss, c->noalloc_ws, c->stream));
    return sync_stream(c);
}

int annf_mlp_adam_dev(annf_ctx *c, float *params, const float *grads, float *m, float *v, long long n, float lr,
                      float beta1, float beta2, float eps, int step) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (n < 0 || step < 1) return fail(ANNF_EINVAL, "bad Adam arguments");
    if (n > 0 && (!params || !grads || !m || !v)) return fail(ANNF_EINVAL, "null pointer");
    RCHK(set_device(c));
    if (n == 0) return ANNF_OK;
    HIPCHK(mlp_adam(params, grads, m, v, n, lr, beta1, beta2, eps, step, c->stream));
    return sync_stream(c);
}

int annf_set_encoder(annf_ctx *c, const float *params, int psize, int h1, int h2, int d) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!params) return fail(ANNF_EINVAL, "null parameters");
    if (!encoder_supported(psize, h1, h2, d))
        return fail(ANNF_EUNSUPPORTED, "encoder p=%d %d-%d-%d: fused inference needs 128-128 hidden layers, d <= 32 "
                    "and a padded patch of at most 512 bytes", psize, h1, h2, d);
    RCHK(set_device(c));
    const int DK = encoder_dk(psize);
    RCHK(c->enc_w1.ensure((size_t)h1 * DK * sizeof(__bf16)));
    RCHK(c->enc_b1.ensure((size_t)h1 * sizeof(float)));
    RCHK(c->enc_w2.ensure((size_t)h2 * h1 * sizeof(__bf16)));
    RCHK(c->enc_b2.ensure((size_t)h2 * sizeof(float)));
    RCHK(c->enc_w3.ensure((size_t)32 * h2 * sizeof(__bf16)));
    RCHK(c->enc_b3.ensure((size_t)32 * sizeof(float)));
    HIPCHK(pack_encoder(params, psize, h1, h2, d, c->enc_w1.as<__bf16>(), c->enc_b1.as<float>(),
                        c->enc_w2.as<__bf16>(), c->enc_b2.as<float>(), c->enc_w3.as<__bf16>(), c->enc_b3.as<float>(),
                        c->stream));
    RCHK(sync_stream(c));
    c->has_encoder = true;
    c->enc_p = psize;
    c->enc_d = d;
    return ANNF_OK;
}

int annf_encode_image_dev(annf_ctx *c, const uint8_t *img, int h, int w, int psize, float *out) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!img || !out) return fail(ANNF_EINVAL, "null pointer");
    if (!c->has_encoder || c->enc_p != psize) return fail(ANNF_EINVAL, "no encoder for patch size %d", psize);
    if (psize < 1 || h < psize || w < psize) return fail(ANNF_EINVAL, "patch size %d larger than the image", psize);
    RCHK(set_device(c));
    const int nc = w - psize + 1;
    const long long n = (long long)(h - psize + 1) * nc;
    const PatchGeom g{img, 1, nc, w, psize, 3 * psize, 3LL * w, 3LL * w * h};
    HIPCHK(encode_patches(g, n, c->enc_w1.as<__bf16>(), c->enc_b1.as<float>(), c->enc_w2.as<__bf16>(),
                          c->enc_b2.as<float>(), c->enc_w3.as<__bf16>(), c->enc_b3.as<float>(), c->enc_d, c->enc_d,
                          out, c->stream));
    return sync_stream(c);
}

int annf_pca_cov_dev(annf_ctx *c, const uint8_t *img_a, int ha, int wa, const int *idx_a, int n_a,
                     const uint8_t *img_b, int hb, int wb, const int *idx_b, int n_b, int psize, double *mean_out,
                     double *cov_out) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!mean_out || !cov_out) return fail(ANNF_EINVAL, "null output");
    if (n_a < 0 || n_b < 0 || n_a + n_b < 2) return fail(ANNF_EINVAL, "need at least two samples");
    if ((n_a > 0 && (!img_a || !idx_a)) || (n_b > 0 && (!img_b || !idx_b))) return fail(ANNF_EINVAL, "null input");
    const int p = psize;
    if (p < 1 || (n_a > 0 && (ha < p || wa < p)) || (n_b > 0 && (hb < p || wb < p)))
        return fail(ANNF_EINVAL, "patch size %d larger than an image", p);
    if ((long long)n_a + n_b > (1LL << 24))
        return fail(ANNF_EUNSUPPORTED, "at most 2^24 samples (exact integer sums in FP64)");
    RCHK(set_device(c));
    const int D = p * p * 3;
    const PatchGeom ga{img_a, 1, n_a > 0 ? wa - p + 1 : 1, wa, p, 3 * p, 3LL * wa, 3LL * wa * ha};
    const PatchGeom gb{img_b, 1, n_b > 0 ? wb - p + 1 : 1, wb, p, 3 * p, 3LL * wb, 3LL * wb * hb};
    RCHK(c->pca_scratch.ensure(pca_cov_scratch(D, n_a + n_b)));
    HIPCHK(pca_cov(ga, idx_a, n_a, gb, idx_b, n_b, D, c->pca_scratch.as<uint8_t>(), mean_out, cov_out, c->stream));
    return sync_stream(c);
}

int annf_pca_fit_dev(annf_ctx *c, const uint8_t *img_a, int ha, int wa, const int *idx_a, int n_a,
                     const uint8_t *img_b, int hb, int wb, const int *idx_b, int n_b, int psize, int d,
                     double *components_out, double *mean_out, double *variance_out) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!components_out) return fail(ANNF_EINVAL, "null output");
    const int D = psize * psize * 3;
    if (d < 1 || d > D) return fail(ANNF_EINVAL, "n_components %d outside [1, %d]", d, D);
    RCHK(set_device(c));
    // covariance into the first D*D doubles of pca_eig, eigenvalues and the
    // off-diagonal workspace after it, then rocSOLVER's status word
    const size_t dd = (size_t)D * D;
    RCHK(c->pca_eig.ensure((dd + 2 * (size_t)D) * sizeof(double) + 256));
    double *C = c->pca_eig.as<double>(), *w = C + dd, *e = w + D;
    rocblas_int *info = reinterpret_cast<rocblas_int *>(e + D);
    RCHK(annf_pca_cov_dev(c, img_a, ha, wa, idx_a, n_a, img_b, hb, wb, idx_b, n_b, psize, mean_out, C));
    if (!c->blas) {
        if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
            c->blas = nullptr;
            return fail(ANNF_EHIP, "rocblas_create_handle failed");
        }
    }
    if (rocblas_set_stream(c->blas, c->stream) != rocblas_status_success)
        return fail(ANNF_EHIP, "rocblas_set_stream failed");
    // C is symmetric, so its row-major layout is also the column-major one
    // rocSOLVER reads; on return its columns are the eigenvectors
    const rocblas_status rs = rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, D, C, D, w, e, info);
    if (rs != rocblas_status_success) return fail(ANNF_EHIP, "rocsolver_dsyevd failed (status %d)", (int)rs);
    HIPCHK(pca_components(C, w, D, d, components_out, variance_out, c->stream));
    rocblas_int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, c->stream));
    RCHK(sync_stream(c));
    if (hinfo != 0) return fail(ANNF_EHIP, "rocsolver_dsyevd did not converge (info %d)", (int)hinfo);
    return ANNF_OK;
}

int annf_patch_fp8_dev(annf_ctx *c, const uint8_t *img, int h, int w, int psize, uint8_t *codes, float *norm,
                       float *inv) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!img) return fail(ANNF_EINVAL, "null image");
    const int p = psize;
    if (p < 1 || h < p || w < p) return fail(ANNF_EINVAL, "patch size %d larger than the image", p);
    if (!leaffp8_supported(p)) return fail(ANNF_EUNSUPPORTED, "patch size %d: the FP8 codes take 1..12", p);
    RCHK(set_device(c));
    const long long n = (long long)(h - p + 1) * (w - p + 1);
    const size_t dq = (size_t)leaffp8_dq(p);
    RCHK(c->f8_qa.ensure((size_t)n * dq));
    RCHK(c->f8_na.ensure((size_t)n * sizeof(float)));
    RCHK(c->f8_ia.ensure((size_t)n * sizeof(float)));
    HIPCHK(patch_fp8(img, h, w, p, c->f8_qa.as<uint8_t>(), c->f8_na.as<float>(), c->f8_ia.as<float>(), c->stream));
    if (codes) HIPCHK(hipMemcpyAsync(codes, c->f8_qa.p, (size_t)n * dq, hipMemcpyDeviceToDevice, c->stream));
    if (norm) HIPCHK(hipMemcpyAsync(norm, c->f8_na.p, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    if (inv) HIPCHK(hipMemcpyAsync(inv, c->f8_ia.p, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    return sync_stream(c);
}

int annf_project_dev(annf_ctx *c, const uint8_t *img, int h, int w, int psize, const void *basis,
                     const double *mean, int d, int projection, float *out) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!img || !out) return fail(ANNF_EINVAL, "null pointer");
    const int p = psize;
    if (p < 1 || h < p || w < p) return fail(ANNF_EINVAL, "patch size %d larger than the image", p);
    if (d < 1 || d > ANNF_MAX_D) return fail(ANNF_EUNSUPPORTED, "dimension %d outside [1, %d]", d, ANNF_MAX_D);
    if (projection != ANNF_PROJ_PCA32 && projection != ANNF_PROJ_PCA64 && projection != ANNF_PROJ_ENCODER)
        return fail(ANNF_EINVAL, "projection %d is not ANNF_PROJ_PCA32, _PCA64 or _ENCODER", projection);
    const bool enc = projection == ANNF_PROJ_ENCODER;
    if (!enc && (!basis || !mean)) return fail(ANNF_EINVAL, "null basis or mean");
    if (enc && (!c->has_encoder || c->enc_p != p || c->enc_d != d))
        return fail(ANNF_EINVAL, "no encoder for patch size %d and dimension %d on this handle", p, d);
    RCHK(set_device(c));
    const int D = p * p * 3, dpad = dpad_of(d), nc = w - p + 1;
    const long long n = (long long)(h - p + 1) * nc;
    const PatchGeom g{img, 1, nc, w, p, 3 * p, 3LL * w, 3LL * w * h};
    RCHK(c->stage_in.ensure((size_t)n * dpad * sizeof(float)));
    float *padded = c->stage_in.as<float>();
    if (enc) {
        HIPCHK(encode_patches(g, n, c->enc_w1.as<__bf16>(), c->enc_b1.as<float>(), c->enc_w2.as<__bf16>(),
                              c->enc_b2.as<float>(), c->enc_w3.as<__bf16>(), c->enc_b3.as<float>(), d, dpad, padded,
                              c->stream, 2));
    } else {
        RCHK(c->basis32.ensure((size_t)D * dpad * sizeof(float)));
        RCHK(c->mean32.ensure((size_t)D * sizeof(float)));
        RCHK(c->basis64.ensure((size_t)D * dpad * sizeof(double)));
        RCHK(c->mean64.ensure((size_t)D * sizeof(double)));
        hipLaunchKernelGGL(k_prep_basis, dim3((D * dpad + 255) / 256), dim3(256), 0, c->stream, basis,
                           (int)(projection == ANNF_PROJ_PCA64), mean, d, D, dpad, c->basis32.as<float>(),
                           c->mean32.as<float>(), c->basis64.as<double>(), c->mean64.as<double>());
        HIPCHK(hipGetLastError());
        if (projection == ANNF_PROJ_PCA32) {
            RCHK(c->proj_pack.ensure(projection_pack_floats(p, dpad) * sizeof(float)));
            HIPCHK(pack_projection(c->basis32.as<float>(), c->mean32.as<float>(), p, dpad, c->proj_pack.as<float>(),
                                   c->stream));
        }
        HIPCHK(project_patches(img, h, w, p, c->basis32.as<float>(), c->mean32.as<float>(), c->basis64.as<double>(),
                               c->mean64.as<double>(), d, dpad, projection == ANNF_PROJ_PCA64, padded,
                               c->proj_pack.as<float>(), 2, c->stream));
    }
    hipLaunchKernelGGL(k_unpad, dim3((unsigned)((n * d + 255) / 256)), dim3(256), 0, c->stream, padded, n, d, dpad,
                       out);
    HIPCHK(hipGetLastError());
    return sync_stream(c);
}

int annf_field_openclopt_dev(annf_ctx *c, const float *red_a, int n_rows, int n_cols, const float *red_b, int n_b,
                             int d, int leaf_size, const uint8_t *img_a, int ha, int wa, const uint8_t *img_b,
                             int hb, int wb, int psize, int *nn_out, int *knn_out, float *knd_out) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!red_a || !red_b) return fail(ANNF_EINVAL, "null reduced patches");
    if (n_rows < 1 || n_cols < 1 || n_b < 1) return fail(ANNF_EINVAL, "empty patch set");
    if ((long long)n_rows * n_cols > 0x7FFFFFFFLL) return fail(ANNF_EINVAL, "too many patches");
    if (d < 1 || d > ANNF_MAX_D) return fail(ANNF_EUNSUPPORTED, "dimension %d outside [1, %d]", d, ANNF_MAX_D);
    if (leaf_size < 1) return fail(ANNF_EINVAL, "leaf_size must be positive");
    const int n_a = n_rows * n_cols;
    // the row chain reads B indices on A's patch grid (driverKNN.fut:250-261)
    if (n_rows > 1 && n_b != n_a)
        return fail(ANNF_EUNSUPPORTED, "propagation needs B on A's patch grid (%d vs %d patches)", n_b, n_a);
    if (nn_out) {
        const int p = psize;
        if (!img_a || !img_b) return fail(ANNF_EINVAL, "the best of the 8 candidates needs both images");
        if (p < 1 || ha < p || wa < p || hb < p || wb < p) return fail(ANNF_EINVAL, "patch size %d larger than an image", p);
        if ((ha - p + 1) != n_rows || (wa - p + 1) != n_cols || (long long)(hb - p + 1) * (wb - p + 1) != n_b)
            return fail(ANNF_EINVAL, "images do not match the patch grids");
    }
    OclTree &t = c->ocl;
    ocl_tree_shape(n_b, leaf_size, t.height, t.q, t.ppl, t.mp);
    if (t.height > 29) return fail(ANNF_EUNSUPPORTED, "tree height %d: the traversal keeps its path in 31 bits",
                                   t.height);
    if ((long long)t.mp * d > 0x7FFFFFFFLL) return fail(ANNF_EUNSUPPORTED, "padded tree too large");
    RCHK(set_device(c));
    c->has_ocl = false;
    const size_t nq = (size_t)std::max(t.q, 1), mp = (size_t)t.mp;
    RCHK(c->ocl_lub.ensure(2 * (size_t)d * sizeof(uint32_t)));
    RCHK(c->ocl_indir.ensure(mp * sizeof(int)));
    RCHK(c->ocl_indir2.ensure(mp * sizeof(int)));
    RCHK(c->ocl_keys.ensure(mp * sizeof(unsigned long long)));
    RCHK(c->ocl_keys2.ensure(mp * sizeof(unsigned long long)));
    RCHK(c->ocl_dims.ensure(nq * sizeof(int)));
    RCHK(c->ocl_clanc.ensure(nq * sizeof(int)));
    RCHK(c->ocl_vals.ensure(nq * sizeof(float)));
    RCHK(c->ocl_leafpts.ensure(mp * d * sizeof(float)));
    RCHK(c->ocl_o2l.ensure(mp * sizeof(int)));
    const size_t sort_bytes = ocl_sort_bytes(t.mp);
    RCHK(c->ocl_sort.ensure(sort_bytes));
    RCHK(c->ocl_nat.ensure((size_t)n_a * sizeof(int)));
    RCHK(c->ocl_pos.ensure((size_t)n_a * 8 * sizeof(int)));
    RCHK(c->ocl_dst.ensure((size_t)n_a * 8 * sizeof(float)));
    t.lub = c->ocl_lub.as<uint32_t>();
    t.indir = c->ocl_indir.as<int>();
    t.indir2 = c->ocl_indir2.as<int>();
    t.keys = c->ocl_keys.as<unsigned long long>();
    t.keys2 = c->ocl_keys2.as<unsigned long long>();
    t.dims = c->ocl_dims.as<int>();
    t.clanc = c->ocl_clanc.as<int>();
    t.vals = c->ocl_vals.as<float>();
    t.leaf_pts = c->ocl_leafpts.as<float>();
    t.orig2leaf = c->ocl_o2l.as<int>();
    HIPCHK(ocl_build(red_b, n_b, d, leaf_size, t, c->ocl_sort.p, sort_bytes, c->stream));
    HIPCHK(ocl_queries(t, red_a, n_rows, n_cols, d, c->ocl_nat.as<int>(), c->ocl_pos.as<int>(),
                       c->ocl_dst.as<float>(), c->stream));
    HIPCHK(ocl_best(c->ocl_pos.as<int>(), t.indir, n_a, img_a, wa, wa - psize + 1, img_b, wb, wb - psize + 1, psize,
                    knn_out, nn_out, c->stream));
    if (knd_out)
        HIPCHK(hipMemcpyAsync(knd_out, c->ocl_dst.p, (size_t)n_a * 8 * sizeof(float), hipMemcpyDeviceToDevice,
                              c->stream));
    RCHK(sync_stream(c));
    c->has_ocl = true;
    return ANNF_OK;
}

int annf_openclopt_tree_export(annf_ctx *c, int *shape, int *dims, float *vals, int *clanc, int *indir,
                               int *orig2leaf) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!c->has_ocl) return fail(ANNF_ENOTREE, "no openclopt-semantics field computed on this handle");
    RCHK(set_device(c));
    const OclTree &t = c->ocl;
    if (shape) { shape[0] = t.height; shape[1] = t.q; shape[2] = t.ppl; shape[3] = t.mp; }
    const size_t q = (size_t)t.q, mp = (size_t)t.mp;
    if (dims && q) HIPCHK(hipMemcpyAsync(dims, t.dims, q * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (vals && q) HIPCHK(hipMemcpyAsync(vals, t.vals, q * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (clanc && q) HIPCHK(hipMemcpyAsync(clanc, t.clanc, q * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (indir) HIPCHK(hipMemcpyAsync(indir, t.indir, mp * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (orig2leaf) HIPCHK(hipMemcpyAsync(orig2leaf, t.orig2leaf, mp * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    return sync_stream(c);
}

int annf_set_timing(annf_ctx *c, int on) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    c->timing = on != 0;
    return ANNF_OK;
}

int annf_stage_times(annf_ctx *c, float *ms, int max_stages) {
    if (!c) return fail(ANNF_EINVAL, "null handle");
    if (!c->have_stages) return 0;
    int n = std::min(max_stages, kStages);
    for (int i = 0; i < n; ++i) ms[i] = c->stage_ms[i];
    return n;
}

}  // extern "C"
