"""CPU oracle for the propagation-assisted k-d-tree NNF path.

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it; the product package ``annfmp_amd`` never does.

Two CPU implementations are exposed through ctypes:

* ``liboracle.so`` -- this repository's C restatement (``annf_oracle.c``) of the
  reference's openmp semantics (annfmp/openmp/kdtree/src/*.c);
* ``_ref/libannf_ref.so`` -- the reference's own openmp C core, compiled from
  its sources by ``oracle/Makefile`` where ``/root/reference`` exists (used to
  pin the restatement and to write the golden fixtures; absent on machines
  that only receive this repository).

Plus numpy restatements of the Python-side stages:

* ``tree_depth``       -- annfmp/openmp/kdtree/base.py:260-300
* ``extract_patches``  -- annfmp/openmp/base.py:152-166 (sklearn extract_patches_2d order)
* ``project_f64``      -- annfmp/openmp/base.py:193-200 (sklearn PCA.transform in FP64, cast to f32)
* ``exact_field``      -- brute-force k-NN over full uint8 patches with exact
                          integer squared distances (equal to FP64: every
                          value is an integer below 2**25)
* ``leaf_knn_exact``   -- the same restricted to a query's probed leaves (the
                          answer of the BF16 full-patch leaf k-NN)
* ``leaf_knn_propagated`` (and ``_rows`` for a sample of the queries) -- the
                          same over the probed leaves plus the leaves the row
                          above's matches propagate (base.c:255-282)
* ``e4m3_rne`` / ``patch_fp8`` -- the FP8 operands of the FP8 full-patch
                          variant (no reference counterpart): E4M3 (the OCP
                          8-bit float format, 4 exponent bits of bias 7, 3
                          mantissa bits, largest 448) by round to nearest
                          even over the enumerated code table; pinned by
                          known codes, ties and round trips in the tests

and the openclopt backend's semantics (``ocl_oracle.c``, restating
annfmp/openclopt/src/futhark/*.fut): ``ocl_tree_shape``, ``OclTree``,
``ocl_brute_force`` / ``ocl_brute_force_par`` and ``ocl_field``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libannf_ref.so")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


# --------------------------------------------------------------------------
# numpy restatements
# --------------------------------------------------------------------------

def tree_depth(n, leaf_size, tree_depth=None):
    """annfmp/openmp/kdtree/base.py:260-300 (math.log float semantics kept)."""
    d = int(math.floor(math.log(n, 2)))
    if tree_depth is not None:
        if tree_depth > d:
            warnings.warn("tree_depth %i too large (using smaller depth): %i.\n" % (tree_depth, d))
            return d
        return tree_depth
    if not leaf_size > 1:
        raise Exception("Parameter 'leaf_size' must be larger than 1")
    subtr = int(math.floor(math.log(leaf_size, 2)))
    if d - subtr < 0:
        warnings.warn("tree_depth smaller than 0; setting tree_depth to 0")
        return 0
    return d - subtr


def extract_patches(img, p):
    """All p x p windows, row-major over the top-left corner, feature
    f = (i*p + j)*3 + ch  (openmp/base.py:152-166)."""
    img = np.ascontiguousarray(img)
    H, W, ch = img.shape
    nr, nc = H - p + 1, W - p + 1
    v = np.lib.stride_tricks.as_strided(
        img, shape=(nr, nc, p, p, ch),
        strides=(img.strides[0], img.strides[1], img.strides[0], img.strides[1], img.strides[2]))
    return np.ascontiguousarray(v).reshape(nr * nc, p * p * ch)


def project_f64(patches, components, mean):
    """sklearn PCA.transform (X @ C.T - mean @ C.T in FP64) then astype(f32)
    (openmp/base.py:195)."""
    comp = np.asarray(components, dtype=np.float64)
    out = np.empty((patches.shape[0], comp.shape[0]), dtype=np.float32)
    off = np.asarray(mean, dtype=np.float64) @ comp.T
    step = 1 << 16
    for s in range(0, patches.shape[0], step):
        x = patches[s:s + step].astype(np.float64)
        out[s:s + step] = (x @ comp.T - off).astype(np.float32)
    return out


def _e4m3_table():
    """Values of the non-negative E4M3 codes 0..126 (127 is NaN), ascending."""
    c = np.arange(127)
    e, m = c >> 3, c & 7
    return np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * 2.0 ** (e.astype(np.float64) - 7))


def e4m3_rne(v):
    """E4M3 codes of float32 values |v| <= 448 (and a rounding's worth above),
    round to nearest, ties to the even code; +0 for zero."""
    v = np.asarray(v, np.float32)
    tab = _e4m3_table()
    a = np.abs(v).astype(np.float64)
    hi = np.clip(np.searchsorted(tab, a), 0, 126)
    lo = np.clip(hi - 1, 0, 126)
    dlo, dhi = a - tab[lo], tab[hi] - a
    pick = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(lo % 2 == 0, lo, hi)))
    pick = np.where(a >= tab[126], 126, pick)
    return (pick | np.where((v < 0) & (pick > 0), 0x80, 0)).astype(np.uint8)


def e4m3_value(code):
    """float64 value of E4M3 codes (no NaN codes expected)."""
    code = np.asarray(code, np.int64)
    return np.where(code & 0x80, -1.0, 1.0) * _e4m3_table()[np.minimum(code & 0x7F, 126)]


def patch_fp8(img, p):
    """Per-patch FP8 codes, norms and inverse scales of every p x p patch:
    codes (n x DQ uint8; row r of the patch at bytes 4 RW r .. with RW =
    ceil(3p / 4) dwords, zero past the row and up to DQ = 64 ceil(4 p RW /
    64)), inv = m / 448 and norm = (sum of squared code values) * inv^2
    (the sum exact in float64, the product in float64, then float32), with
    m = max |x - 128| over the patch (inv = 1 and scale 1 when m = 0)."""
    pa = extract_patches(img, p).astype(np.int64)           # n x 3 p^2, feature (r * p + c) * 3 + ch
    n = pa.shape[0]
    rl = 3 * p
    rw = (rl + 3) // 4
    dq = (p * rw * 4 + 63) // 64 * 64
    x = pa.reshape(n, p, rl) - 128
    m = np.abs(x).reshape(n, -1).max(axis=1)
    m32 = m.astype(np.float32)
    scale = np.where(m > 0, np.float32(448.0) / np.maximum(m32, np.float32(1)), np.float32(1)).astype(np.float32)
    inv = np.where(m > 0, m32 / np.float32(448.0), np.float32(1)).astype(np.float32)
    v = (x.astype(np.float32) * scale[:, None, None]).astype(np.float32)
    codes_rows = e4m3_rne(v)                                  # n x p x 3p
    codes = np.zeros((n, dq), np.uint8)
    for r in range(p):
        codes[:, 4 * rw * r:4 * rw * r + rl] = codes_rows[:, r, :]
    vals = e4m3_value(codes_rows).reshape(n, -1)
    ssum = (vals * vals).sum(axis=1)
    norm = (ssum * (inv.astype(np.float64) * inv.astype(np.float64))).astype(np.float32)
    return codes, norm, inv


def exact_field(pa, pb, k, chunk=2048):
    """Exact k nearest B patches of every A patch by squared L2 on uint8
    values.  Integer arithmetic (int64), ties broken by lower index.
    Returns (idx [n_a, k] int64, sqdist [n_a, k] int64)."""
    a = pa.astype(np.int64)
    b = pb.astype(np.int64)
    bn = (b * b).sum(1)
    bt = b.T.copy()
    idx = np.empty((a.shape[0], k), np.int64)
    dst = np.empty((a.shape[0], k), np.int64)
    for s in range(0, a.shape[0], chunk):
        x = a[s:s + chunk]
        d2 = (x * x).sum(1)[:, None] + bn[None, :] - 2 * (x @ bt)
        part = np.argpartition(d2, kth=min(k, d2.shape[1] - 1) - 1 if k < d2.shape[1] else d2.shape[1] - 1, axis=1)[:, :k]
        # exact tie order: sort candidates by (distance, index) among all
        # entries not greater than the k-th distance
        kth = np.take_along_axis(d2, part, 1).max(1)
        for r in range(x.shape[0]):
            cand = np.nonzero(d2[r] <= kth[r])[0]
            order = np.lexsort((cand, d2[r, cand]))[:k]
            idx[s + r] = cand[order]
            dst[s + r] = d2[r, cand[order]]
    return idx, dst


def exact_pair_d2(pa, pb, ia, ib):
    """Exact integer squared L2 between patch rows pa[ia] and pb[ib]."""
    x = pa[np.asarray(ia, np.int64)].astype(np.int64)
    y = pb[np.asarray(ib, np.int64)].astype(np.int64)
    return ((x - y) ** 2).sum(1)


def leaf_knn_exact(pa, pb, qleaf, leaves, orig, nprobe, k):
    """Exact k nearest B patches of every A patch among the points of the
    leaves qleaf ^ 0 .. qleaf ^ (nprobe - 1) (the full-patch leaf k-NN of
    the precision variants).  qleaf: natural leaf per A patch; leaves:
    [from, to) tree positions per leaf; orig: B index per tree position.
    Integer squared L2, ties by lower B index; missing entries are -1.
    Returns (idx [n_a, k] int64, sqdist [n_a, k] int64).

    No reference counterpart: the reference scans leaves in the projected
    space (kdtree.c:305-352).  With nprobe = number of leaves it is the
    exact field (exact_field)."""
    a = pa.astype(np.int64)
    b = pb.astype(np.int64)
    qleaf = np.asarray(qleaf, np.int64)
    idx = np.full((a.shape[0], k), -1, np.int64)
    dst = np.full((a.shape[0], k), -1, np.int64)
    for leaf in np.unique(qleaf):
        rows = np.nonzero(qleaf == leaf)[0]
        pos = np.concatenate([np.arange(leaves[2 * (leaf ^ r)], leaves[2 * (leaf ^ r) + 1])
                              for r in range(nprobe)])
        cand = np.asarray(orig, np.int64)[pos]
        x = a[rows]
        y = b[cand]
        d2 = (x * x).sum(1)[:, None] + (y * y).sum(1)[None, :] - 2 * (x @ y.T)
        order = np.lexsort((np.broadcast_to(cand, d2.shape), d2), axis=1)[:, :k]
        m = order.shape[1]
        idx[rows, :m] = cand[order]
        dst[rows, :m] = np.take_along_axis(d2, order, 1)
    return idx, dst


def leaf_knn_propagated(pa, pb, qleaf, leaves, orig, pli, n_cols, k, nprobe=1, prop_k=None, chain=False):
    """The full-patch leaf k-NN with the parallel propagated pass: pass one is
    leaf_knn_exact over the leaves qleaf ^ 0 .. qleaf ^ (nprobe - 1); a patch
    below the first row then also scans the leaves pli[o'] of the first
    prop_k (default k) pass-one matches o of its row-above neighbour, shifted
    one row down (o' = o + n_cols when that is a B patch, base.c:265-276),
    leaving out the leaves pass one scanned and repeats, and keeps the k
    smallest (squared distance, B index) over all scanned points.  chain:
    the matches of the row above are its final ones (rows in order, the
    reference's row sweep, base.c:227-306) instead of its pass-one ones.
    Returns (idx [n_a, k], sqdist [n_a, k]) int64, -1 past the candidates."""
    prop_k = k if prop_k is None else min(prop_k, k)
    i1, d1 = leaf_knn_exact(pa, pb, qleaf, leaves, orig, nprobe, k)
    n_a, n_b = pa.shape[0], pb.shape[0]
    out_i, out_d = i1.copy(), d1.copy()
    A = pa.astype(np.int64)
    B = pb.astype(np.int64)
    orig = np.asarray(orig, np.int64)
    for q in range(n_cols, n_a):
        ls = [int(qleaf[q]) ^ r for r in range(nprobe)]
        above = out_i if chain else i1
        for j in range(prop_k):
            o = int(above[q - n_cols, j])
            if o < 0:
                continue
            if o + n_cols < n_b:
                o += n_cols
            lf = int(pli[o])
            if lf not in ls:
                ls.append(lf)
        if len(ls) == nprobe:
            continue
        cand = orig[np.concatenate([np.arange(leaves[2 * lf], leaves[2 * lf + 1]) for lf in ls])]
        d2 = ((A[q][None, :] - B[cand]) ** 2).sum(1)
        order = np.lexsort((cand, d2))[:k]
        out_i[q, :order.size] = cand[order]
        out_d[q, :order.size] = d2[order]
    return out_i, out_d


def leaf_knn_propagated_rows(pa, pb, qleaf, leaves, orig, pli, n_cols, k, rows, nprobe=1, prop_k=None):
    """leaf_knn_propagated for the A patches listed in `rows` only (for
    workloads too large to scan every query in Python): pass one is computed
    for those rows and for the patches directly above them, then each row
    scans its probed and propagated leaves as leaf_knn_propagated does.
    qleaf: natural leaf of every A patch.  Returns (idx, sqdist) of shape
    [len(rows), k], int64, -1 past the candidates."""
    rows = np.asarray(rows, np.int64)
    prop_k = k if prop_k is None else min(prop_k, k)
    n_b = pb.shape[0]
    above = np.where(rows >= n_cols, rows - n_cols, rows)
    both = np.unique(np.concatenate([rows, above]))
    i1, d1 = leaf_knn_exact(pa[both], pb, np.asarray(qleaf)[both], leaves, orig, nprobe, k)
    at = {int(r): j for j, r in enumerate(both)}
    out_i = np.full((rows.size, k), -1, np.int64)
    out_d = np.full((rows.size, k), -1, np.int64)
    orig = np.asarray(orig, np.int64)
    for j, q in enumerate(rows.tolist()):
        out_i[j], out_d[j] = i1[at[q]], d1[at[q]]
        if q < n_cols:
            continue
        ls = [int(qleaf[q]) ^ r for r in range(nprobe)]
        for m in range(prop_k):
            o = int(i1[at[q - n_cols], m])
            if o < 0:
                continue
            if o + n_cols < n_b:
                o += n_cols
            lf = int(pli[o])
            if lf not in ls:
                ls.append(lf)
        if len(ls) == nprobe:
            continue
        cand = orig[np.concatenate([np.arange(leaves[2 * lf], leaves[2 * lf + 1]) for lf in ls])]
        d2 = ((pa[q].astype(np.int64)[None, :] - pb[cand].astype(np.int64)) ** 2).sum(1)
        order = np.lexsort((cand, d2))[:k]
        out_i[j, :order.size] = cand[order]
        out_d[j, :order.size] = d2[order]
    return out_i, out_d


# --------------------------------------------------------------------------
# the C restatement (liboracle.so)
# --------------------------------------------------------------------------

class _OrcTree(C.Structure):
    _fields_ = [("n", C.c_int), ("d", C.c_int), ("depth", C.c_int),
                ("pts", _fp), ("orig", _ip), ("axis", _ip), ("split", _fp),
                ("leaves", _ip), ("pli", _ip)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = C.CDLL(ORACLE_SO)
        L.orc_build_tree.argtypes = [_fp, C.c_int, C.c_int, C.c_int, C.POINTER(_OrcTree)]
        L.orc_build_tree.restype = C.c_int
        L.orc_natural_leaf.argtypes = [C.POINTER(_OrcTree), _fp]
        L.orc_natural_leaf.restype = C.c_int
        L.orc_exact_knn.argtypes = [C.POINTER(_OrcTree), _fp, C.c_int, C.c_long, _fp, _ip]
        L.orc_scan_leaves.argtypes = [C.POINTER(_OrcTree), _fp, _ip, C.c_int, C.c_int, _fp, _ip]
        L.orc_neighbors.argtypes = [C.POINTER(_OrcTree), _fp, C.c_int, C.c_int, C.c_long, _fp, _ip, C.c_int]
        L.orc_neighbors.restype = C.c_int
        L.orc_norm_array.argtypes = [_u8p, _u8p, C.c_int]
        L.orc_norm_array.restype = C.c_float
        L.orc_process_rows.argtypes = [C.POINTER(_OrcTree), _u8p, _u8p, C.c_int, _fp, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _fp, C.c_int]
        L.orc_process_rows.restype = C.c_int
        L.orc_rows_range.argtypes = [C.POINTER(_OrcTree), _fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, _ip, _fp, C.c_int]
        L.orc_rows_range.restype = C.c_int
        L.orc_best_of_k_images.argtypes = [_u8p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                           _ip, _fp, C.c_int]
        L.orc_best_of_k_images.restype = C.c_int
        i = C.c_int
        L.ocl_tree_shape.argtypes = [i, i, _ip, _ip, _ip, _ip]
        L.ocl_build_tree.argtypes = [_fp, i, i, i, _ip, _fp, _ip, _ip, _fp, _ip]
        L.ocl_find_leaf.argtypes = [_ip, _fp, i, _fp]
        L.ocl_find_leaf.restype = i
        L.ocl_brute_force.argtypes = [_fp, i, _fp, i, i, _ip, _fp]
        L.ocl_brute_force_par.argtypes = [_fp, i, _fp, i, i, _ip, _fp]
        L.ocl_exact_knn.argtypes = [i, _ip, _fp, _ip, _fp, i, i, _fp, i, _ip, _fp]
        L.ocl_field.argtypes = [_fp, i, i, _fp, i, i, i, _u8p, _u8p, i, _ip, _ip, _fp, _ip, _ip, _fp, _ip, _ip, _ip,
                                i]
        _lib = L
    return _lib


class OracleTree:
    """A k-d tree held in numpy arrays (tree positions = XI order)."""

    def __init__(self, X, depth):
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, d = X.shape
        self.n, self.d, self.depth = n, d, depth
        self.pts = np.empty((n, d), np.float32)
        self.orig = np.empty(n, np.int32)
        self.axis = np.zeros(max((1 << depth) - 1, 1), np.int32)
        self.split = np.zeros(max((1 << depth) - 1, 1), np.float32)
        self.leaves = np.empty(2 << depth, np.int32)
        self.pli = np.empty(n, np.int32)
        self._s = _OrcTree(n, d, depth, _ptr(self.pts, _fp), _ptr(self.orig, _ip), _ptr(self.axis, _ip),
                           _ptr(self.split, _fp), _ptr(self.leaves, _ip), _ptr(self.pli, _ip))
        rc = lib().orc_build_tree(_ptr(X, _fp), n, d, depth, C.byref(self._s))
        if rc != 0:
            raise ValueError("orc_build_tree failed")
        self.n_nodes = (1 << depth) - 1
        self.axis = self.axis[:self.n_nodes]
        self.split = self.split[:self.n_nodes]

    def natural_leaf(self, q):
        q = np.ascontiguousarray(q, dtype=np.float32)
        return lib().orc_natural_leaf(C.byref(self._s), _ptr(q, _fp))

    def natural_leaves(self, Q):
        return np.array([self.natural_leaf(q) for q in np.asarray(Q, np.float32)], np.int32)

    def exact_knn(self, q, K, max_leaves=2147483640):
        q = np.ascontiguousarray(q, dtype=np.float32)
        ld = np.empty(K, np.float32)
        lp = np.empty(K, np.int32)
        lib().orc_exact_knn(C.byref(self._s), _ptr(q, _fp), K, max_leaves, _ptr(ld, _fp), _ptr(lp, _ip))
        return ld, lp

    def scan_leaves(self, q, leaf_list, K):
        q = np.ascontiguousarray(q, dtype=np.float32)
        ll = np.ascontiguousarray(leaf_list, dtype=np.int32)
        ld = np.empty(K, np.float32)
        lp = np.empty(K, np.int32)
        lib().orc_scan_leaves(C.byref(self._s), _ptr(q, _fp), _ptr(ll, _ip), len(ll), K,
                              _ptr(ld, _fp), _ptr(lp, _ip))
        return ld, lp

    def neighbors(self, Q, K, max_leaves=2147483640, nthreads=0):
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        dist = np.empty((Q.shape[0], K), np.float32)
        idx = np.empty((Q.shape[0], K), np.int32)
        rc = lib().orc_neighbors(C.byref(self._s), _ptr(Q, _fp), Q.shape[0], K, max_leaves,
                                 _ptr(dist, _fp), _ptr(idx, _ip), nthreads)
        if rc != 0:
            raise ValueError("orc_neighbors failed")
        return dist, idx

    def process_rows(self, red, n_rows, n_cols, K, propagate=True, select_best_nn=True,
                     pa=None, pb=None, nthreads=0):
        """Returns (final [N] int32, lists_pos [N,K] tree positions, lists_dst [N,K])."""
        red = np.ascontiguousarray(red, dtype=np.float32)
        N = red.shape[0]
        if select_best_nn:
            pa = np.ascontiguousarray(pa, dtype=np.uint8)
            pb = np.ascontiguousarray(pb, dtype=np.uint8)
            D = pa.shape[1]
        else:
            pa = pb = np.zeros((1, 1), np.uint8)
            D = 1
        final = np.empty(N, np.int32)
        pos = np.empty((N, K), np.int32)
        dst = np.empty((N, K), np.float32)
        rc = lib().orc_process_rows(C.byref(self._s), _ptr(pa, _u8p), _ptr(pb, _u8p), D, _ptr(red, _fp), N,
                                    n_rows, n_cols, K, int(bool(propagate)), int(bool(select_best_nn)),
                                    _ptr(final, _ip), _ptr(pos, _ip), _ptr(dst, _fp), nthreads)
        if rc != 0:
            raise ValueError("orc_process_rows failed")
        return final, pos, dst

    def rows_range(self, red, n_cols, K, propagate, row_begin, row_end, pos, dst, nthreads=0):
        red = np.ascontiguousarray(red, dtype=np.float32)
        return lib().orc_rows_range(C.byref(self._s), _ptr(red, _fp), red.shape[0], n_cols, K,
                                    int(bool(propagate)), row_begin, row_end, _ptr(pos, _ip), _ptr(dst, _fp),
                                    nthreads)


def norm_array(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    return lib().orc_norm_array(_ptr(a, _u8p), _ptr(b, _u8p), a.size)


def best_of_k_images(img_a, img_b, p, cand, nthreads=0):
    """Best-of-K over original B indices cand [N, K], reading windows of the
    two images directly.  Returns (final int32 [N], dist float32 [N])."""
    img_a = np.ascontiguousarray(img_a, dtype=np.uint8)
    img_b = np.ascontiguousarray(img_b, dtype=np.uint8)
    H, W, _ = img_a.shape
    nr, nc = H - p + 1, W - p + 1
    cand = np.ascontiguousarray(cand, dtype=np.int32)
    final = np.empty(nr * nc, np.int32)
    fd = np.empty(nr * nc, np.float32)
    lib().orc_best_of_k_images(_ptr(img_a, _u8p), _ptr(img_b, _u8p), W, p, nr, nc, _ptr(cand, _ip),
                               cand.shape[1], _ptr(final, _ip), _ptr(fd, _fp), nthreads)
    return final, fd


# --------------------------------------------------------------------------
# the reference's own C core (oracle/_ref/libannf_ref.so), container only
# --------------------------------------------------------------------------

class _RefParams(C.Structure):          # global.h:18-27
    _fields_ = [("n_neighbors", C.c_int), ("tree_depth", C.c_int), ("max_leaves", C.c_int),
                ("num_threads", C.c_int), ("verbosity_level", C.c_int), ("splitting_type", C.c_int)]


class _RefNode(C.Structure):            # global.h:30-35
    _fields_ = [("axis", C.c_int), ("splitting_value", C.c_float)]


class _RefRecord(C.Structure):          # global.h:38-50
    _fields_ = [("Xtrain", _fp), ("nXtrain", C.c_int), ("dXtrain", C.c_int), ("XI", C.c_void_p),
                ("nodes", C.POINTER(_RefNode)), ("leaves", _ip), ("points_leaf_indices", _ip),
                ("tree_depth", C.c_int)]


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference core not built (oracle/_ref/libannf_ref.so)")
        L = C.CDLL(REF_SO)
        L.init_extern.argtypes = [C.c_int] * 6 + [C.POINTER(_RefParams)]
        L.fit_extern.argtypes = [_fp, C.c_int, C.c_int, C.POINTER(_RefRecord), C.POINTER(_RefParams)]
        L.neighbors_extern.argtypes = [_fp, C.c_int, C.c_int, _fp, C.c_int, C.c_int, _ip, C.c_int, C.c_int,
                                       C.POINTER(_RefRecord), C.POINTER(_RefParams)]
        L.process_rows_extern.argtypes = [_ip, C.c_int, _u8p, C.c_int, C.c_int, _u8p, C.c_int, C.c_int,
                                          _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.POINTER(_RefRecord), C.POINTER(_RefParams)]
        L.brute_force_group.argtypes = [_fp, C.c_int, _ip, C.c_int, C.POINTER(_RefRecord), C.POINTER(_RefParams),
                                        _ip, _fp, C.c_int]
        L.brute_force_group.restype = C.c_int
        L.norm_array.argtypes = [_u8p, _u8p, C.c_int]
        L.norm_array.restype = C.c_float
        L.kd_tree_query_tree_sequential.argtypes = [_fp, _fp, _ip, C.c_int, C.c_int, C.c_int,
                                                    C.POINTER(_RefRecord)]
        L.kd_tree_query_tree_sequential.restype = C.c_int
        _ref = L
    return _ref


class RefTree:
    """The reference's own tree (fit_extern), with arrays copied out."""

    def __init__(self, X, depth, K=1, max_leaves=2147483640, num_threads=1):
        L = ref_lib()
        self.X = np.ascontiguousarray(X, dtype=np.float32)   # borrowed by the record
        n, d = self.X.shape
        self.n, self.d, self.depth = n, d, depth
        self.params = _RefParams()
        # init_extern sets the process-wide OpenMP thread count (base.c:41)
        L.init_extern(K, depth, max_leaves, num_threads, 1, 0, C.byref(self.params))
        self.rec = _RefRecord()
        L.fit_extern(_ptr(self.X, _fp), n, d, C.byref(self.rec), C.byref(self.params))
        nn = (1 << depth) - 1
        raw = np.ctypeslib.as_array(C.cast(self.rec.XI, C.POINTER(C.c_uint8)), shape=(n * (4 * d + 4),))
        rec = raw.reshape(n, 4 * d + 4)
        self.pts = rec[:, :4 * d].copy().view(np.float32).reshape(n, d)
        self.orig = rec[:, 4 * d:].copy().view(np.int32).reshape(n)
        if nn > 0:
            nodes = np.ctypeslib.as_array(C.cast(self.rec.nodes, C.POINTER(C.c_int32)), shape=(2 * nn,))
            nodes = nodes.reshape(nn, 2).copy()
            self.axis = nodes[:, 0].copy()
            self.split = nodes[:, 1].copy().view(np.float32)
        else:
            self.axis = np.zeros(0, np.int32)
            self.split = np.zeros(0, np.float32)
        self.leaves = np.ctypeslib.as_array(self.rec.leaves, shape=(2 << depth,)).copy()
        self.pli = np.ctypeslib.as_array(self.rec.points_leaf_indices, shape=(n,)).copy()

    def set_k(self, K):
        self.params.n_neighbors = K

    def neighbors(self, Q, K):
        self.set_k(K)
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        dist = np.zeros((Q.shape[0], K), np.float32)
        idx = np.zeros((Q.shape[0], K), np.int32)
        ref_lib().neighbors_extern(_ptr(Q, _fp), Q.shape[0], Q.shape[1], _ptr(dist, _fp), Q.shape[0], K,
                                   _ptr(idx, _ip), Q.shape[0], K, C.byref(self.rec), C.byref(self.params))
        return dist, idx

    def brute_force_group(self, q, leaf_list, K, propagate=1):
        q = np.ascontiguousarray(q, dtype=np.float32)
        ll = np.ascontiguousarray(leaf_list, dtype=np.int32)
        idx = np.zeros(K, np.int32)
        dist = np.zeros(K, np.float32)
        ref_lib().brute_force_group(_ptr(q, _fp), q.size, _ptr(ll, _ip), K, C.byref(self.rec),
                                    C.byref(self.params), _ptr(idx, _ip), _ptr(dist, _fp), int(propagate))
        return dist, idx

    def natural_leaf(self, q, K=1):
        q = np.ascontiguousarray(q, dtype=np.float32)
        dist = np.zeros(K, np.float32)
        idx = np.zeros(K, np.int32)
        return ref_lib().kd_tree_query_tree_sequential(_ptr(q, _fp), _ptr(dist, _fp), _ptr(idx, _ip), K, 1, 1,
                                                       C.byref(self.rec))

    def exact_knn(self, q, K):
        q = np.ascontiguousarray(q, dtype=np.float32)
        dist = np.zeros(K, np.float32)
        idx = np.zeros(K, np.int32)
        ref_lib().kd_tree_query_tree_sequential(_ptr(q, _fp), _ptr(dist, _fp), _ptr(idx, _ip), K, 2147483640, 0,
                                                C.byref(self.rec))
        return dist, idx

    def process_rows(self, pa, pb, red, n_rows, n_cols, K, propagate=True, select_best_nn=True):
        pa = np.ascontiguousarray(pa, dtype=np.uint8)
        pb = np.ascontiguousarray(pb, dtype=np.uint8)
        red = np.ascontiguousarray(red, dtype=np.float32)
        final = np.zeros(red.shape[0], np.int32)
        ref_lib().process_rows_extern(_ptr(final, _ip), final.size, _ptr(pa, _u8p), pa.shape[0], pa.shape[1],
                                      _ptr(pb, _u8p), pb.shape[0], pb.shape[1], _ptr(red, _fp), red.shape[0],
                                      red.shape[1], n_rows, n_cols, K, int(bool(propagate)),
                                      int(bool(select_best_nn)), C.byref(self.rec), C.byref(self.params))
        return final


def ref_norm_array(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.uint8)
    return ref_lib().norm_array(_ptr(a, _u8p), _ptr(b, _u8p), a.size)


# --------------------------------------------------------------------------
# the openclopt backend's semantics (ocl_oracle.c)
# --------------------------------------------------------------------------

OCL_K = 8


def ocl_tree_shape(m, leaf_size):
    """buildKDtree.fut:19-30: (height, internal nodes, points per leaf, padded count)."""
    h, q, ppl, mp = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    lib().ocl_tree_shape(int(m), int(leaf_size), C.byref(h), C.byref(q), C.byref(ppl), C.byref(mp))
    return h.value, q.value, ppl.value, mp.value


class OclTree:
    """The openclopt tree of X (m x d float32): dims / vals / clanc per
    internal node, indir (tree position -> padded point), leaf_pts (rows in
    tree order, +inf pads) and orig2leaf."""

    def __init__(self, X, leaf_size):
        X = np.ascontiguousarray(X, dtype=np.float32)
        m, d = X.shape
        self.m, self.d = m, d
        self.height, self.q, self.ppl, self.mp = ocl_tree_shape(m, leaf_size)
        nq = max(self.q, 1)
        self.dims = np.zeros(nq, np.int32)
        self.vals = np.zeros(nq, np.float32)
        self.clanc = np.zeros(nq, np.int32)
        self.indir = np.empty(self.mp, np.int32)
        self.leaf_pts = np.empty((self.mp, d), np.float32)
        self.orig2leaf = np.empty(self.mp, np.int32)
        lib().ocl_build_tree(_ptr(X, _fp), m, d, int(leaf_size), _ptr(self.dims, _ip), _ptr(self.vals, _fp),
                             _ptr(self.clanc, _ip), _ptr(self.indir, _ip), _ptr(self.leaf_pts, _fp),
                             _ptr(self.orig2leaf, _ip))
        self.dims, self.vals, self.clanc = self.dims[:self.q], self.vals[:self.q], self.clanc[:self.q]

    def find_leaf(self, qv):
        qv = np.ascontiguousarray(qv, dtype=np.float32)
        dims = self.dims if self.q else np.zeros(1, np.int32)
        vals = self.vals if self.q else np.zeros(1, np.float32)
        return lib().ocl_find_leaf(_ptr(dims, _ip), _ptr(vals, _fp), self.height, _ptr(qv, _fp))

    def leaf(self, l):
        return self.leaf_pts[l * self.ppl:(l + 1) * self.ppl]

    def natural_list(self, qv):
        """findNaturalLeaves for one query: (positions, squared distances)."""
        ki = np.full(OCL_K, -1, np.int32)
        kd = np.full(OCL_K, np.finfo(np.float32).max, np.float32)
        l = self.find_leaf(qv)
        ocl_brute_force(qv, self.leaf(l), l * self.ppl, ki, kd)
        return ki, kd

    def exact_knn(self, qv):
        """exactKnn for one query from its natural-leaf list."""
        qv = np.ascontiguousarray(qv, dtype=np.float32)
        ki, kd = self.natural_list(qv)
        dims = self.dims if self.q else np.zeros(1, np.int32)
        vals = self.vals if self.q else np.zeros(1, np.float32)
        clanc = self.clanc if self.q else np.full(1, -1, np.int32)
        lib().ocl_exact_knn(self.height, _ptr(dims, _ip), _ptr(vals, _fp), _ptr(clanc, _ip),
                            _ptr(self.leaf_pts, _fp), self.d, self.ppl, _ptr(qv, _fp), self.find_leaf(qv),
                            _ptr(ki, _ip), _ptr(kd, _fp))
        return ki, kd


def ocl_brute_force(qv, pts, beg, ki, kd):
    """brute-force.fut:5-22 on (ki, kd) in place."""
    qv = np.ascontiguousarray(qv, dtype=np.float32)
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    lib().ocl_brute_force(_ptr(qv, _fp), qv.size, _ptr(pts, _fp), int(beg), pts.shape[0], _ptr(ki, _ip),
                          _ptr(kd, _fp))


def ocl_brute_force_par(qv, pts, beg, ki, kd):
    """brute-force.fut:25-66 on (ki, kd) in place."""
    qv = np.ascontiguousarray(qv, dtype=np.float32)
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    lib().ocl_brute_force_par(_ptr(qv, _fp), qv.size, _ptr(pts, _fp), int(beg), pts.shape[0], _ptr(ki, _ip),
                              _ptr(kd, _fp))


def ocl_field(red_a, n_rows, n_cols, red_b, leaf_size, pa=None, pb=None, rows=0):
    """The openclopt pipeline on reduced patches: dict of nn (when pa/pb
    given), knn / knd (n_a x 8: original indices, squared reduced
    distances), nat (natural leaves) and the tree arrays.  rows > 0: only
    the first `rows` rows (outputs cut to them)."""
    red_a = np.ascontiguousarray(red_a, dtype=np.float32)
    red_b = np.ascontiguousarray(red_b, dtype=np.float32)
    n_a, d = red_a.shape
    m = red_b.shape[0]
    assert n_a == n_rows * n_cols and red_b.shape[1] == d
    h, q, ppl, mp = ocl_tree_shape(m, leaf_size)
    out = {"knn": np.empty((n_a, OCL_K), np.int32), "knd": np.empty((n_a, OCL_K), np.float32),
           "nat": np.empty(n_a, np.int32), "dims": np.zeros(max(q, 1), np.int32),
           "vals": np.zeros(max(q, 1), np.float32), "clanc": np.zeros(max(q, 1), np.int32),
           "indir": np.empty(mp, np.int32), "orig2leaf": np.empty(mp, np.int32)}
    nn = np.full(n_a, -1, np.int32)
    D = 0
    if pa is not None:
        pa = np.ascontiguousarray(pa, dtype=np.uint8)
        pb = np.ascontiguousarray(pb, dtype=np.uint8)
        D = pa.shape[1]
    lib().ocl_field(_ptr(red_a, _fp), n_rows, n_cols, _ptr(red_b, _fp), m, d, int(leaf_size),
                    _ptr(pa, _u8p) if pa is not None else None, _ptr(pb, _u8p) if pb is not None else None, D,
                    _ptr(nn, _ip), _ptr(out["knn"], _ip), _ptr(out["knd"], _fp), _ptr(out["nat"], _ip),
                    _ptr(out["dims"], _ip), _ptr(out["vals"], _fp), _ptr(out["clanc"], _ip), _ptr(out["indir"], _ip),
                    _ptr(out["orig2leaf"], _ip), int(rows))
    if 0 < rows < n_rows:
        for k in ("knn", "knd", "nat"):
            out[k] = out[k][:rows * n_cols]
        nn = nn[:rows * n_cols]
    for k in ("dims", "vals", "clanc"):
        out[k] = out[k][:q]
    if pa is not None:
        out["nn"] = nn
    out["shape"] = (h, q, ppl, mp)
    return out
