"""The CPU oracle against the golden vectors written by the reference's own
openmp C core (tests/golden/make_golden.py), and against closed forms."""
import math

import numpy as np
import pytest

from conftest import golden, TREE_FIXTURES, ROW_FIXTURES
from oracle import oracle as O


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name", TREE_FIXTURES)
def test_tree_matches_reference(name):
    g = golden(name)
    t = O.OracleTree(g["X"], int(g["depth"]))
    assert np.array_equal(bits(t.pts), bits(g["pts"]))
    assert np.array_equal(t.orig, g["orig"])
    assert np.array_equal(t.axis, g["axis"])
    assert np.array_equal(bits(t.split), bits(g["split"]))
    assert np.array_equal(t.leaves, g["leaves"])
    assert np.array_equal(t.pli, g["pli"])


@pytest.mark.parametrize("name", TREE_FIXTURES)
def test_queries_match_reference(name):
    g = golden(name)
    t = O.OracleTree(g["X"], int(g["depth"]))
    K = int(g["K"])
    for i, q in enumerate(g["Q"]):
        assert t.natural_leaf(q) == g["natural"][i]
        d, p = t.exact_knn(q, K)
        assert np.array_equal(p, g["exact_pos"][i]) and np.array_equal(bits(d), bits(g["exact_dist"][i]))
        d, p = t.scan_leaves(q, g["group_leaves"][i], K)
        assert np.array_equal(p, g["group_pos"][i]) and np.array_equal(bits(d), bits(g["group_dist"][i]))
    d, idx = t.neighbors(g["Q"], K, nthreads=1)
    assert np.array_equal(idx, g["neighbors_idx"])
    assert np.array_equal(bits(d), bits(g["neighbors_dist"]))


@pytest.mark.parametrize("name", ROW_FIXTURES)
def test_process_rows_matches_reference(name):
    g = golden(name)
    p = int(g["psize"])
    pa = O.extract_patches(g["image_a"], p)
    pb = O.extract_patches(g["image_b"], p)
    t = O.OracleTree(g["red_b"], int(g["depth"]))
    final, pos, dst = t.process_rows(g["red_a"], int(g["n_rows"]), int(g["n_cols"]), int(g["K"]),
                                     bool(g["propagate"]), bool(g["select_best_nn"]), pa, pb, nthreads=2)
    assert np.array_equal(final, g["final"])
    if "final_k5" in g:
        f5, _, _ = t.process_rows(g["red_a"], int(g["n_rows"]), int(g["n_cols"]), 5, bool(g["propagate"]),
                                  bool(g["select_best_nn"]), pa, pb, nthreads=2)
        assert np.array_equal(f5, g["final_k5"])


def test_threads_do_not_change_results():
    g = golden("rows_k3.npz")
    p = int(g["psize"])
    pa = O.extract_patches(g["image_a"], p)
    pb = O.extract_patches(g["image_b"], p)
    t = O.OracleTree(g["red_b"], int(g["depth"]))
    a = t.process_rows(g["red_a"], int(g["n_rows"]), int(g["n_cols"]), 3, True, True, pa, pb, nthreads=1)
    b = t.process_rows(g["red_a"], int(g["n_rows"]), int(g["n_cols"]), 3, True, True, pa, pb, nthreads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_norm_array_matches_reference():
    g = golden("norm_array.npz")
    out = np.array([O.norm_array(a, b) for a, b in zip(g["A"], g["B"])], np.float32)
    assert np.array_equal(bits(out), bits(g["out"]))
    # the all-0 vs all-255 pair: 363 * 255^2 = 23,604,075 needs more than 24
    # bits, so the sequential FP32 sum (23,603,972) is not the exact integer;
    # the reference's value is sqrt of the FP32 sum, not of the integer
    s = np.float32(0)
    for _ in range(363):
        s = np.float32(s + np.float32(255 * 255))
    assert float(s) == 23603972.0
    assert g["out"][0] == np.float32(math.sqrt(float(s)))
    assert g["out"][0] != np.float32(math.sqrt(363 * 255 * 255))


def test_patch_order_closed_form():
    """extract_patches: feature f = (i*p + j)*3 + ch of window (y, x)."""
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (9, 11, 3)).astype(np.uint8)
    p = 4
    P = O.extract_patches(img, p)
    nc = 11 - p + 1
    for n in (0, 5, 17, P.shape[0] - 1):
        y, x = divmod(n, nc)
        for f in (0, 7, 20, 47):
            i, rem = divmod(f, p * 3)
            j, ch = divmod(rem, 3)
            assert P[n, f] == img[y + i, x + j, ch]


def test_tree_depth_closed_form():
    for n in (2, 3, 100, 1023, 1024, 1025, 62500, 2043700, 8246464):
        for leaf in (2, 5, 32, 64, 128):
            want = max(0, int(math.floor(math.log(n, 2))) - int(math.floor(math.log(leaf, 2))))
            with np.errstate(all="ignore"):
                import warnings
                with warnings.catch_warnings():
                    warnings.simplefilter("ignore")
                    assert O.tree_depth(n, leaf) == want
    assert O.tree_depth(2043700, 64) == 14
    assert O.tree_depth(62500, 32) == 10
    assert O.tree_depth(8246464, 128) == 15


def _seq_dist(P, q):
    """Sequential FP32 sum of (p - q)^2 in numpy float32 arithmetic."""
    s = np.zeros(P.shape[0], np.float32)
    for j in range(P.shape[1]):
        t = (P[:, j] - q[j]).astype(np.float32)
        s = (s + (t * t).astype(np.float32)).astype(np.float32)
    return s


def test_exact_search_equals_brute_force():
    """Row-0 search is exact: same list as a stable sort of all points."""
    rng = np.random.default_rng(7)
    X = rng.standard_normal((3000, 6)).astype(np.float32)
    t = O.OracleTree(X, O.tree_depth(3000, 16))
    for q in rng.standard_normal((30, 6)).astype(np.float32):
        d, p = t.exact_knn(q, 5)
        dd = _seq_dist(t.pts, q)
        order = np.lexsort((np.arange(len(dd)), dd))[:5]
        assert np.array_equal(bits(d), bits(dd[order]))
        assert np.array_equal(p, order)


def test_partition_invariants():
    """Every internal node: left holds all x < split and right all x > split;
    leaf sizes differ by at most one; pli inverts the leaf ranges."""
    rng = np.random.default_rng(3)
    X = rng.integers(0, 5, (2000, 3)).astype(np.float32)
    depth = O.tree_depth(2000, 8)
    t = O.OracleTree(X, depth)
    sizes = np.diff(t.leaves.reshape(-1, 2), axis=1).ravel()
    assert sizes.max() - sizes.min() <= 1 or depth == 0
    for l in range(1 << depth):
        a, b = t.leaves[2 * l], t.leaves[2 * l + 1]
        assert np.all(t.pli[t.orig[a:b]] == l)
    # walk: every point's natural descent with its own coordinates reaches a
    # leaf whose path splits bound it
    bnd = [(0, 2000)]
    for lvl in range(depth):
        nb = []
        for j, (lo, hi) in enumerate(bnd):
            node = (1 << lvl) - 1 + j
            c = hi - lo
            ax, v = t.axis[node], t.split[node]
            left = t.pts[lo:lo + c // 2, ax]
            right = t.pts[lo + c // 2:hi, ax]
            assert np.all(left <= v) and np.all(right >= v)
            # split = value of rank c//2 of the node's axis column (util.c Wirth select)
            assert np.sort(t.pts[lo:hi, ax])[c // 2] == v
            nb += [(lo, lo + c // 2), (lo + c // 2, hi)]
        bnd = nb


def test_planted_mistake_is_caught():
    """A tie rule flipped in a Python copy of the insert (<= -> <) changes a
    golden list: the fixtures pin the rule."""
    g = golden("tree_ties.npz")
    t = O.OracleTree(g["X"], int(g["depth"]))
    K = int(g["K"])

    def scan(q, leaves, strict):
        ld = [np.float32(3.402823466e38)] * K
        lp = [0] * K
        for l in leaves:
            for i in range(t.leaves[2 * l], t.leaves[2 * l + 1]):
                d = _seq_dist(t.pts[i:i + 1], q)[0]
                rej = (ld[-1] < d) if strict else (ld[-1] <= d)
                if rej:
                    continue
                ld[-1], lp[-1] = d, i
                j = K - 1
                while j > 0 and ld[j] < ld[j - 1]:
                    ld[j], ld[j - 1] = ld[j - 1], ld[j]
                    lp[j], lp[j - 1] = lp[j - 1], lp[j]
                    j -= 1
        return np.array(lp)

    ok = [np.array_equal(scan(q, g["group_leaves"][i], False), g["group_pos"][i]) for i, q in enumerate(g["Q"][:16])]
    bad = [np.array_equal(scan(q, g["group_leaves"][i], True), g["group_pos"][i]) for i, q in enumerate(g["Q"][:16])]
    assert all(ok)
    assert not all(bad)


def test_leaf_knn_exact_reduces_to_exact_field():
    """Probing every leaf makes the leaf-restricted k-NN the exact field;
    one leaf restricts it to that leaf's points (brute force by hand)."""
    rng = np.random.default_rng(11)
    pa = rng.integers(0, 256, (300, 27)).astype(np.uint8)
    pb = rng.integers(0, 256, (256, 27)).astype(np.uint8)
    pb[5] = pb[9]                                   # an exact tie between two B patches
    pa[0] = pb[9]
    leaves = np.array([0, 64, 64, 128, 128, 192, 192, 256])
    orig = rng.permutation(256)
    qleaf = rng.integers(0, 4, 300)
    i1, d1 = O.leaf_knn_exact(pa, pb, qleaf, leaves, orig, 4, 6)
    i2, d2 = O.exact_field(pa, pb, 6)
    assert np.array_equal(i1, i2) and np.array_equal(d1, d2)
    assert i1[0, 0] == 5 and d1[0, 0] == 0 and i1[0, 1] == 9     # tie by lower index
    i3, d3 = O.leaf_knn_exact(pa, pb, qleaf, leaves, orig, 1, 3)
    for r in range(0, 300, 37):
        l = qleaf[r]
        cand = orig[leaves[2 * l]:leaves[2 * l + 1]]
        dd = [(int(((pa[r].astype(int) - pb[c].astype(int)) ** 2).sum()), int(c)) for c in cand]
        dd.sort()
        assert [c for _, c in dd[:3]] == list(i3[r]) and [x for x, _ in dd[:3]] == list(d3[r])


@pytest.mark.parametrize("nprobe,prop_k", [(1, None), (2, None), (1, 2), (2, 1)])
def test_leaf_knn_propagated_is_the_best_of_its_leaves(nprobe, prop_k):
    """Every query's list is exactly ordered, lists no B index twice, and
    nothing among the points of its probed leaves and its propagated leaves
    (collected here row by row, as base.c:255-282 does, from the first
    prop_k matches above) is closer than its k-th entry."""
    rng = np.random.default_rng(5)
    n_rows, n_cols, D, k = 9, 11, 12, 3
    n = n_rows * n_cols
    pa = rng.integers(0, 256, (n, D)).astype(np.uint8)
    pb = rng.integers(0, 256, (n, D)).astype(np.uint8)
    leaves = np.array([0, 25, 25, 50, 50, 74, 74, 99])
    orig = rng.permutation(n)
    pli = np.empty(n, np.int64)
    for lf in range(4):
        pli[orig[leaves[2 * lf]:leaves[2 * lf + 1]]] = lf
    qleaf = rng.integers(0, 4, n)
    i1, _ = O.leaf_knn_exact(pa, pb, qleaf, leaves, orig, nprobe, k)
    ip, dp = O.leaf_knn_propagated(pa, pb, qleaf, leaves, orig, pli, n_cols, k, nprobe=nprobe, prop_k=prop_k)
    pk = k if prop_k is None else prop_k
    for q in range(n):
        allowed = {int(qleaf[q]) ^ r for r in range(nprobe)}
        if q >= n_cols:
            for o in i1[q - n_cols][:pk]:
                o2 = o + n_cols if o + n_cols < n else o
                allowed.add(int(pli[o2]))
        pts = np.concatenate([orig[leaves[2 * lf]:leaves[2 * lf + 1]] for lf in sorted(allowed)])
        d_all = sorted((int(((pa[q].astype(int) - pb[j].astype(int)) ** 2).sum()), int(j)) for j in pts)
        assert [j for _, j in d_all[:k]] == list(ip[q]) and [x for x, _ in d_all[:k]] == list(dp[q])
        assert len(set(ip[q].tolist())) == k
    assert (ip[:n_cols] == i1[:n_cols]).all()


@pytest.mark.parametrize("nprobe,prop_k", [(1, None), (2, 2)])
def test_leaf_knn_chain_is_the_best_of_its_leaves(nprobe, prop_k):
    """With chain=True a query's propagated leaves come from the final list
    of the patch above (rows in order, base.c:227-306): every list is the
    exact best over its probed leaves and those leaves, collected here from
    the function's own output one row up; and the chain reaches points the
    parallel form does not."""
    rng = np.random.default_rng(9)
    n_rows, n_cols, D, k = 10, 9, 12, 3
    n = n_rows * n_cols
    pa = rng.integers(0, 256, (n, D)).astype(np.uint8)
    pb = rng.integers(0, 256, (n, D)).astype(np.uint8)
    leaves = np.array([0, 12, 12, 23, 23, 34, 34, 45, 45, 56, 56, 67, 67, 78, 78, 90])
    orig = rng.permutation(n)
    pli = np.empty(n, np.int64)
    for lf in range(8):
        pli[orig[leaves[2 * lf]:leaves[2 * lf + 1]]] = lf
    qleaf = rng.integers(0, 8, n)
    ic, dc = O.leaf_knn_propagated(pa, pb, qleaf, leaves, orig, pli, n_cols, k, nprobe=nprobe, prop_k=prop_k,
                                   chain=True)
    ip, _ = O.leaf_knn_propagated(pa, pb, qleaf, leaves, orig, pli, n_cols, k, nprobe=nprobe, prop_k=prop_k)
    pk = k if prop_k is None else prop_k
    for q in range(n):
        allowed = {int(qleaf[q]) ^ r for r in range(nprobe)}
        if q >= n_cols:
            for o in ic[q - n_cols][:pk]:
                o2 = o + n_cols if o + n_cols < n else o
                allowed.add(int(pli[o2]))
        pts = np.concatenate([orig[leaves[2 * lf]:leaves[2 * lf + 1]] for lf in sorted(allowed)])
        d_all = sorted((int(((pa[q].astype(int) - pb[j].astype(int)) ** 2).sum()), int(j)) for j in pts)
        assert [j for _, j in d_all[:k]] == list(ic[q]) and [x for x, _ in d_all[:k]] == list(dc[q])
    assert (ic[:2 * n_cols] == ip[:2 * n_cols]).all()     # rows 0 and 1 agree by construction
    assert (ic != ip).any()


@pytest.mark.parametrize("shape,p,d,n_fit", [((256, 256), 7, 10, 10000), ((300, 310), 11, 16, 5000),
                                              ((64, 80), 5, 40, 2000)])
def test_project_f64_is_sklearns_transform(shape, p, d, n_fit):
    """oracle.project_f64 equals scikit-learn's PCA.transform followed by
    astype(float32) (openmp/base.py:193-200) bit for bit, on a basis fitted
    the reference's way (extract_patches_2d samples of both images,
    openmp/base.py:168-191); the 300x310 case crosses the oracle's 65,536-row
    block boundary.  The projection of the fitted mean patch is zero and the
    map is linear, which a dropped mean term or a transposed basis breaks."""
    from sklearn.decomposition import PCA
    from sklearn.feature_extraction.image import extract_patches_2d
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.small_spec(*shape), seed=p)
    sa = extract_patches_2d(a, (p, p), max_patches=n_fit // 2, random_state=1).reshape(n_fit // 2, -1)
    sb = extract_patches_2d(b, (p, p), max_patches=n_fit // 2, random_state=2).reshape(n_fit // 2, -1)
    pca = PCA(n_components=d).fit(np.concatenate([sa, sb]))
    pa = O.extract_patches(a, p)
    assert pa.shape[0] > (1 << 16) or shape[0] < 300
    got = O.project_f64(pa, pca.components_, pca.mean_)
    want = pca.transform(pa).astype(np.float32)
    assert got.dtype == np.float32 and got.shape == (pa.shape[0], d)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # closed forms: the mean maps to 0; linearity on integer combinations
    z = O.project_f64(pca.mean_[None, :], pca.components_, pca.mean_)
    assert np.abs(z).max() <= 1e-9 * np.abs(pca.components_).sum()
    comp = pca.components_.astype(np.float64)
    lin = (pa[:50].astype(np.float64) - pca.mean_) @ comp.T
    assert np.allclose(got[:50], lin, rtol=1e-6, atol=1e-4)


@pytest.mark.parametrize("nprobe,prop_k", [(1, None), (2, 2)])
def test_leaf_knn_propagated_rows_equals_the_full_pass(nprobe, prop_k):
    """The sampled form gives the full form's rows (first row included)."""
    rng = np.random.default_rng(8)
    n_rows, n_cols, D, k = 12, 13, 10, 4
    n = n_rows * n_cols
    pa = rng.integers(0, 256, (n, D)).astype(np.uint8)
    pb = rng.integers(0, 256, (n, D)).astype(np.uint8)
    leaves = np.array([0, 40, 40, 78, 78, 117, 117, 156])
    orig = rng.permutation(n)
    pli = np.empty(n, np.int64)
    for lf in range(4):
        pli[orig[leaves[2 * lf]:leaves[2 * lf + 1]]] = lf
    qleaf = rng.integers(0, 4, n)
    fi, fd = O.leaf_knn_propagated(pa, pb, qleaf, leaves, orig, pli, n_cols, k, nprobe=nprobe, prop_k=prop_k)
    rows = np.array([0, 5, 12, 13, 40, 77, 155])
    si, sd = O.leaf_knn_propagated_rows(pa, pb, qleaf, leaves, orig, pli, n_cols, k, rows, nprobe=nprobe,
                                        prop_k=prop_k)
    assert np.array_equal(si, fi[rows]) and np.array_equal(sd, fd[rows])


def test_e4m3_rounding_known_codes():
    """The FP8 operands' E4M3 codes (OCP 8-bit float, bias 7, 3 mantissa
    bits, largest 448): known codes, ties to the even code, subnormals, and
    every code's value rounds back to itself."""
    v = np.float32([0.0, 1.0, 448.0, 1.0625, 1.1875, -3.5, 2.0 ** -9, 2.0 ** -10, 3 * 2.0 ** -11, 240.0,
                    447.99, 0.875, 17.0, 18.0])
    want = [0x00, 0x38, 0x7E, 0x38, 0x3A, 0x80 | 0x46, 0x01, 0x00, 0x01, 0x77, 0x7E, 0x36, 0x58, 0x59]
    assert O.e4m3_rne(v).tolist() == want
    codes = np.arange(127)
    vals = O.e4m3_value(codes).astype(np.float32)
    assert np.array_equal(O.e4m3_rne(vals), codes.astype(np.uint8))
    assert np.array_equal(O.e4m3_rne(-vals[1:]), (codes[1:] | 0x80).astype(np.uint8))
    # midpoints go to the even neighbour, points just off them to the nearer one
    mid = ((vals[1:-1].astype(np.float64) + vals[2:]) / 2).astype(np.float32)
    got = O.e4m3_rne(mid).astype(np.int64)
    lo = codes[1:-1]
    assert np.array_equal(got, np.where(lo % 2 == 0, lo, lo + 1))
    assert np.array_equal(O.e4m3_rne(np.nextafter(mid, np.float32(0))), lo.astype(np.uint8))


def test_patch_fp8_definition():
    """oracle.patch_fp8 on a hand-made image: scale from the patch's largest
    |x - 128|, rows padded to whole dwords, flat patches coded as zeros with
    inverse scale 1, the norm of the de-scaled codes."""
    img = np.full((4, 5, 3), 128, np.uint8)
    img[0, 0] = (128 + 64, 128 - 32, 128 + 1)
    img[3, 4] = (0, 255, 128)
    codes, norm, inv = O.patch_fp8(img, 2)
    assert codes.shape == (12, 64)              # 6-byte rows -> 2 dwords, 2 rows -> 16 bytes -> 64
    # patch 0 holds img[0:2, 0:2]: m = 64, scale 7, values 448, -224, 7
    assert inv[0] == np.float32(64 / 448)
    assert codes[0, :3].tolist() == [0x7E, O.e4m3_rne(np.float32(-224.0)).item(), O.e4m3_rne(np.float32(7.0)).item()]
    assert (codes[0, 3:].astype(int) == 0).all()
    want = (448.0 ** 2 + 224.0 ** 2 + 7.0 ** 2) * (np.float64(inv[0]) ** 2)
    assert norm[0] == np.float32(want)
    # a flat patch
    assert inv[5] == 1.0 and norm[5] == 0.0 and (codes[5] == 0).all()
    # the last patch holds img[2:4, 3:5]: m = 128 (the 0 byte), scale 3.5
    last = codes[-1]
    assert inv[-1] == np.float32(128 / 448)
    assert last[8 + 3:8 + 6].tolist() == [O.e4m3_rne(np.float32(-448.0)).item(),
                                         O.e4m3_rne(np.float32(127 * 3.5)).item(), 0]
