"""GPU parity: libannf_hip.so against the golden vectors of the reference's
openmp core and against the CPU oracle on the same float32 inputs.

The bar is bit-exact: tree arrays (axis, split bits, leaf ranges, tree
order, leaf of every point), k-lists and index fields.
"""
import numpy as np
import pytest

from conftest import golden, TREE_FIXTURES, ROW_FIXTURES
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _fit(X, depth, K=1, splitting="longest"):
    from annfmp_amd import KDTree
    t = KDTree(n_neighbors=K, leaf_size=2, tree_depth=depth, float_type="float", splitting_type=splitting)
    return t.fit(X)


def _assert_tree_equal(arr, ref):
    assert np.array_equal(arr["axis"], ref.axis)
    assert np.array_equal(bits(arr["split"]), bits(ref.split))
    assert np.array_equal(arr["leaves"], ref.leaves)
    assert np.array_equal(arr["xi_orig"], ref.orig)
    assert np.array_equal(bits(arr["xi_pts"]), bits(ref.pts))
    assert np.array_equal(arr["pli"], ref.pli)


@pytest.mark.parametrize("name", TREE_FIXTURES)
def test_tree_and_neighbors_match_reference(name, gpu_available):
    g = golden(name)
    K = int(g["K"])
    t = _fit(g["X"], int(g["depth"]), K)
    arr = t.tree_arrays()
    assert np.array_equal(arr["axis"], g["axis"])
    assert np.array_equal(bits(arr["split"]), bits(g["split"]))
    assert np.array_equal(arr["leaves"], g["leaves"])
    assert np.array_equal(arr["xi_orig"], g["orig"])
    assert np.array_equal(bits(arr["xi_pts"]), bits(g["pts"]))
    assert np.array_equal(arr["pli"], g["pli"])
    d, idx = t.kneighbors(g["Q"])
    assert np.array_equal(idx, g["neighbors_idx"])
    assert np.array_equal(bits(d), bits(g["neighbors_dist"]))


@pytest.mark.parametrize("name", ROW_FIXTURES)
def test_process_rows_matches_reference(name, gpu_available):
    g = golden(name)
    p = int(g["psize"])
    pa = O.extract_patches(g["image_a"], p)
    pb = O.extract_patches(g["image_b"], p)
    Ks = [int(g["K"])] + ([5] if "final_k5" in g else [])
    for K, want in zip(Ks, [g["final"], g.get("final_k5")]):
        t = _fit(g["red_b"], int(g["depth"]), K)
        nn = np.zeros(g["red_a"].shape[0], np.int32)
        t.process_rows(nn, pa, pb, g["red_a"], int(g["n_rows"]), int(g["n_cols"]), K,
                       bool(g["propagate"]), bool(g["select_best_nn"]))
        assert np.array_equal(nn, want), (name, K, int((nn != want).sum()))


@pytest.mark.parametrize("n,d,kind,leaf", [
    (100003, 16, "gauss", 64),
    (65536, 10, "ties", 32),
    (40000, 3, "dups", 16),
    (5000, 32, "gauss", 8),
    (30001, 40, "gauss", 16),
    (20000, 64, "ties", 16),
    (12000, 47, "dups", 8),
])
def test_tree_matches_oracle(n, d, kind, leaf, gpu_available):
    rng = np.random.default_rng(n + d)
    if kind == "gauss":
        X = rng.standard_normal((n, d)).astype(np.float32)
    elif kind == "ties":
        X = rng.integers(-2, 3, (n, d)).astype(np.float32)
    else:
        X = rng.standard_normal((n, d)).astype(np.float32)
        X[: n // 4] = X[n // 4: 2 * (n // 4)]
        X[:, 1] = 0.25
    depth = O.tree_depth(n, leaf)
    ref = O.OracleTree(X, depth)
    _assert_tree_equal(_fit(X, depth).tree_arrays(), ref)


def test_refits_of_other_sizes_on_one_handle(gpu_available):
    """One handle builds trees of different point counts and depths in turn
    (the level bounds it keeps between builds must follow), with a tree
    import of another size in between."""
    from annfmp_amd import KDTree, _lib
    from annfmp_amd._lib import check, lib, ptr
    rng = np.random.default_rng(11)
    sizes = [(30000, 6, 9), (70001, 6, 11), (30000, 6, 9), (30000, 6, 10)]
    X0 = rng.standard_normal((5000, 6)).astype(np.float32)
    donor = KDTree(n_neighbors=1, leaf_size=2, tree_depth=7, float_type="float", splitting_type="longest").fit(X0)
    arr0 = donor.tree_arrays()
    t = KDTree(n_neighbors=1, leaf_size=2, tree_depth=9, float_type="float", splitting_type="longest")
    for it, (n, d, depth) in enumerate(sizes):
        X = rng.standard_normal((n, d)).astype(np.float32)
        t.tree_depth = depth
        t.fit(X)
        _assert_tree_equal(t.tree_arrays(), O.OracleTree(X, depth))
        if it == 1:
            # an import of a 5000-point, depth-7 tree on this handle
            h = t._init_native(7)
            check("annf_tree_import", lib().annf_tree_import(
                h.h, 5000, 6, 7, ptr(arr0["axis"], _lib._ip), ptr(arr0["split"], _lib._fp),
                ptr(arr0["leaves"], _lib._ip), ptr(arr0["xi_orig"], _lib._ip), ptr(arr0["xi_pts"], _lib._fp),
                ptr(arr0["pli"], _lib._ip)))
            assert np.array_equal(t.tree_arrays()["leaves"], arr0["leaves"])


def test_cyclic_split_matches_oracle_rule(gpu_available):
    """splitting_type='cyclic' uses axis = depth % d (kdtree.c:369-371)."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((4096, 5)).astype(np.float32)
    t = _fit(X, 6, splitting="cyclic")
    arr = t.tree_arrays()
    for lvl in range(6):
        assert np.all(arr["axis"][(1 << lvl) - 1:(1 << (lvl + 1)) - 1] == lvl % 5)


@pytest.mark.parametrize("K,d", [(1, 8), (3, 8), (8, 8), (12, 8), (8, 40), (5, 64), (40, 8), (64, 12)])
def test_exact_search_matches_oracle(K, d, gpu_available):
    rng = np.random.default_rng(K + d)
    X = rng.standard_normal((20000, d)).astype(np.float32)
    Q = rng.standard_normal((500, d)).astype(np.float32)
    depth = O.tree_depth(20000, 32)
    ref = O.OracleTree(X, depth)
    t = _fit(X, depth, K)
    d, idx = t.kneighbors(Q)
    rd, ridx = ref.neighbors(Q, K)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(bits(d), bits(rd))


@pytest.mark.parametrize("K,d,shift", [(16, 6, 40.0), (32, 3, 8.0), (5, 16, 0.0)])
def test_exact_search_far_queries_match_oracle(K, d, shift, gpu_available):
    """Queries far outside clustered data: almost every leaf passes the
    natural leaf's radius, so the queue of candidate leaves fills and is
    bounded, sorted and scanned several times per query."""
    rng = np.random.default_rng(70 + K)
    centres = rng.standard_normal((12, d)).astype(np.float32) * 4
    X = (centres[rng.integers(0, 12, 24000)] + rng.standard_normal((24000, d)).astype(np.float32) * 0.5)
    X = X.astype(np.float32)
    Q = (rng.standard_normal((300, d)) * 3 + shift).astype(np.float32)
    depth = O.tree_depth(24000, 8)
    ref = O.OracleTree(X, depth)
    t = _fit(X, depth, K)
    dd, idx = t.kneighbors(Q)
    rd, ridx = ref.neighbors(Q, K)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(bits(dd), bits(rd))


def test_max_leaves_limits_the_search(gpu_available):
    from annfmp_amd import KDTree
    rng = np.random.default_rng(9)
    X = rng.standard_normal((8192, 4)).astype(np.float32)
    Q = rng.standard_normal((64, 4)).astype(np.float32)
    depth = O.tree_depth(8192, 16)
    ref = O.OracleTree(X, depth)
    t = KDTree(n_neighbors=3, leaf_size=16, max_leaves=2, float_type="float", splitting_type="longest").fit(X)
    d, idx = t.kneighbors(Q)
    rd, ridx = ref.neighbors(Q, 3, max_leaves=2)
    assert np.array_equal(idx, ridx)


def _field_vs_oracle(a, b, p, d, leaf, K, propagate=True, select_best_nn=True, projection="fp64"):
    from annfmp_amd import field_on_device
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    X = np.concatenate([pa[::7], pb[::7]]).astype(np.float64)
    mu = X.mean(0)
    w, V = np.linalg.eigh(np.cov(X.T))
    comp = V[:, ::-1][:, :d].T.copy()
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    depth = O.tree_depth(pb.shape[0], leaf)
    res = field_on_device(a, b, p, comp, mu, depth, K, propagate, select_best_nn, projection=projection,
                          return_reduced=True)
    # projection against the FP64 restatement of sklearn's transform
    ra = O.project_f64(pa, comp, mu)
    scale = np.abs(ra).max()
    tol = 1e-6 if projection == "fp64" else 2e-5
    assert np.abs(res["red_a"] - ra).max() <= tol * scale
    # rows: the oracle on the GPU's own float32 reduced arrays
    ref = O.OracleTree(res["red_b"], depth)
    final, pos, _ = ref.process_rows(res["red_a"], nr, nc, K, propagate, select_best_nn, pa, pb)
    assert np.array_equal(res["knn"], ref.orig[pos])
    assert np.array_equal(res["nn"], final)
    # distance field: norm_array of the chosen match
    chosen = np.array([O.norm_array(pa[i], pb[final[i]]) for i in range(0, len(final), 97)], np.float32)
    assert np.array_equal(bits(res["dist"][::97]), bits(chosen))
    return res


@pytest.mark.parametrize("K,propagate,select", [(1, True, True), (5, True, True), (3, False, True),
                                                (4, True, False), (40, True, True), (64, True, True)])
def test_field_256_matches_oracle(K, propagate, select, gpu_available):
    from annfmp_amd import synth
    a, b = synth.make_pair("256x256-shift", seed=0)
    _field_vs_oracle(a, b, 7, 10, 32, K, propagate, select)


@pytest.mark.parametrize("p,d", [(5, 12), (9, 24), (13, 8), (7, 40), (5, 64)])
def test_field_fp32_projection(p, d, gpu_available):
    """FP32 projection: the MFMA GEMM with one (d <= 16) and two (d > 16)
    component tiles, and the VALU kernel for p > 12 or d > 32."""
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=4)
    _field_vs_oracle(a, b, p, d, 16, 3, projection="fp32")


@pytest.mark.parametrize("d", [16, 32, 40])
def test_field_example_photographs_match_oracle(d, gpu_available):
    """Real photographs: a 160 x 240 crop of the reference's example pair,
    at its example's K = 8, p = 8, leaf 128 (examples/simple.py:16-28),
    with its dim_reduced = 40 and two smaller dimensions."""
    g = golden("example_crop.npz")
    _field_vs_oracle(g["a"], g["b"], 8, d, 128, 8)


def test_field_1080p_matches_oracle(gpu_available):
    """The 1920x1080 workload (p=11, d=16, leaf 64, K=5), full size."""
    from annfmp_amd import synth
    a, b = synth.make_pair("1080p-warp", seed=0)
    _field_vs_oracle(a, b, 11, 16, 64, 5, projection="fp32")


@pytest.mark.parametrize("K", [2, 8, 20, 48])
def test_exact_search_ties_match_oracle(K, gpu_available):
    """Many equal distances across leaves: integer points with duplicates and
    integer queries; the tie order must be the reference's visit order."""
    rng = np.random.default_rng(40 + K)
    X = rng.integers(-3, 4, (30000, 3)).astype(np.float32)
    Q = rng.integers(-3, 4, (400, 3)).astype(np.float32)
    depth = O.tree_depth(30000, 16)
    ref = O.OracleTree(X, depth)
    t = _fit(X, depth, K)
    d, idx = t.kneighbors(Q)
    rd, ridx = ref.neighbors(Q, K)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(bits(d), bits(rd))


def test_process_rows_flat_regions_match_oracle(gpu_available):
    """A pair with large flat areas (many queries share a natural leaf and
    many distances tie) through the reference-shaped process_rows entry."""
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.PairSpec(160, 200, 12, 1.0, grid=(4, 4), max_offset=8), seed=9)
    a[20:120, 30:150] = 77          # one large noiseless block in A
    p, d, leaf, K = 5, 6, 16, 5
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    X = np.concatenate([pa[::5], pb[::5]]).astype(np.float64)
    mu = X.mean(0)
    _, V = np.linalg.eigh(np.cov(X.T))
    comp = V[:, ::-1][:, :d].T.copy()
    ra = O.project_f64(pa, comp, mu)
    rb = O.project_f64(pb, comp, mu)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    depth = O.tree_depth(rb.shape[0], leaf)
    ref = O.OracleTree(rb, depth)
    want, _, _ = ref.process_rows(ra, nr, nc, K, True, True, pa, pb)
    t = _fit(rb, depth, K)
    nn = np.zeros(ra.shape[0], np.int32)
    t.process_rows(nn, pa, pb, ra, nr, nc, K, True, True)
    assert np.array_equal(nn, want)


@pytest.mark.parametrize("k", [1, 5, 12])
def test_exact_field_matches_brute_force(k, gpu_available):
    """annf_exact_field_dev against the integer brute force of the oracle."""
    from annfmp_amd import synth, metrics
    a, b = synth.make_pair(synth.small_spec(48, 64), seed=3)
    p = 5
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    q = metrics.sample_queries(pa.shape[0], 300, seed=k)
    idx, d2 = metrics.exact_knn(a, b, p, q, k)
    ridx, rd2 = O.exact_field(pa[q], pb, k)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(d2, rd2)


def test_exact_field_p11_matches_brute_force(gpu_available):
    """363-byte patches (p = 11, the 1080p workload's patch) on a smaller pair."""
    from annfmp_amd import synth, metrics
    a, b = synth.make_pair(synth.small_spec(140, 180), seed=8)
    p = 11
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    q = metrics.sample_queries(pa.shape[0], 200, seed=2)
    idx, d2 = metrics.exact_knn(a, b, p, q, 8)
    ridx, rd2 = O.exact_field(pa[q], pb, 8)
    assert np.array_equal(idx, ridx) and np.array_equal(d2, rd2)


@pytest.mark.parametrize("p,k,shape", [(7, 32, (200, 210)), (12, 9, (60, 80)), (3, 5, (150, 260))])
def test_exact_field_splits_ties_and_patch_sizes(p, k, shape, gpu_available):
    """The int8-MFMA exact field over several B ranges (more than 32,768 B
    patches), with flat regions (many equal distances: ties go to the lower
    index), k up to 32 and patch sizes of 1 to 14 k-steps, against the
    integer brute force; 333 queries (not a multiple of the block's 128)."""
    from annfmp_amd import synth, metrics
    a, b = synth.make_pair(synth.small_spec(*shape), seed=p + k)
    a[10:40, 20:70] = 90
    b[30:60, 100:160] = 90
    b[70:90, 5:45] = 91
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    q = metrics.sample_queries(pa.shape[0], 333, seed=k)
    q[:20] = np.arange(20) + (12 * (shape[1] - p + 1) + 25)    # queries inside the flat block
    idx, d2 = metrics.exact_knn(a, b, p, q, k)
    ridx, rd2 = O.exact_field(pa[q], pb, k)
    assert np.array_equal(d2, rd2)
    assert np.array_equal(idx, ridx)


def test_field_high_dimension_many_neighbours_match_oracle(gpu_available):
    """d = 48 and K = 64 together (the widest lists and points of the build):
    tree, row 0, propagation sweep and best-of-K against the oracle."""
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.small_spec(90, 120), seed=31)
    _field_vs_oracle(a, b, 6, 48, 128, 64)
