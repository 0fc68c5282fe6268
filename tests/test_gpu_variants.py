"""GPU tests of the precision variants (annf_field_mm_dev): full-patch leaf
k-NN on MFMA.

BF16 distances are exact integers, so the K list must equal the oracle's
exact k-NN over the probed leaves bit for bit (leaf_knn_exact, itself pinned
to the brute-force exact field).  FP8 lists are checked by properties: every
entry is a probed-leaf point with its exact distance order, and the list
agrees with the exact one on most queries.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _basis(pa, pb, d):
    X = np.concatenate([pa[::7], pb[::7]]).astype(np.float64)
    mu = X.mean(0)
    _, V = np.linalg.eigh(np.cov(X.T))
    return V[:, ::-1][:, :d].T.copy(), mu


def _run(a, b, p, d, leaf, K, nprobe, precision, shortlist=20, rows=None):
    from annfmp_amd import field_mm_on_device
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    comp, mu = _basis(pa, pb, d)
    depth = O.tree_depth(pb.shape[0], leaf)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=nprobe, precision=precision,
                             shortlist=shortlist, return_reduced=True)
    # the GPU tree is bit-exact to the oracle's tree over the same reduced B
    tree = O.OracleTree(res["red_b"], depth)
    rows = np.arange(pa.shape[0]) if rows is None else rows
    qleaf = tree.natural_leaves(res["red_a"][rows])
    want_i, want_d = O.leaf_knn_exact(pa[rows], pb, qleaf, tree.leaves, tree.orig, nprobe, K)
    return res, pa, pb, tree, rows, qleaf, want_i, want_d


def _check_bf16(res, rows, want_i, want_d):
    assert np.array_equal(res["knn"][rows], want_i.astype(np.int32))
    assert np.array_equal(res["nn"][rows], want_i[:, 0].astype(np.int32))
    expect = np.sqrt(want_d[:, 0].astype(np.float64)).astype(np.float32)
    assert np.array_equal(res["dist"][rows].view(np.uint32), expect.view(np.uint32))


@pytest.mark.parametrize("p,K,nprobe", [(5, 1, 1), (5, 5, 2), (7, 8, 1), (8, 3, 4), (9, 5, 1), (11, 5, 2)])
def test_bf16_matches_exact_leaf_knn(p, K, nprobe, gpu_available):
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=p)
    res, pa, pb, tree, rows, qleaf, wi, wd = _run(a, b, p, 8, 32, K, nprobe, "bf16")
    _check_bf16(res, rows, wi, wd)


def test_bf16_multi_tile_leaves(gpu_available):
    """Leaves of 256+ points take several 128-point tiles per block."""
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.small_spec(120, 160), seed=21)
    res, pa, pb, tree, rows, qleaf, wi, wd = _run(a, b, 7, 6, 300, 6, 2, "bf16")
    assert (np.diff(tree.leaves.reshape(-1, 2), axis=1) > 128).all()
    _check_bf16(res, rows, wi, wd)


def test_bf16_flat_regions_and_crowded_leaves(gpu_available):
    """A large noiseless block in A sends thousands of queries to one leaf
    (many 64-query chunks of one leaf) and creates equal distances."""
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.PairSpec(160, 200, 12, 1.0, grid=(4, 4), max_offset=8), seed=9)
    a[20:120, 30:150] = 77
    b[60:90, 60:120] = 77
    res, pa, pb, tree, rows, qleaf, wi, wd = _run(a, b, 5, 6, 16, 5, 1, "bf16")
    assert np.bincount(qleaf).max() > 1000
    _check_bf16(res, rows, wi, wd)


def test_bf16_1080p_sampled_queries(gpu_available):
    """The 1920x1080 workload (p=11, PCA-16, leaf 64, K=5), full size; the
    oracle checks the queries of 400 sampled leaves."""
    from annfmp_amd import synth
    from annfmp_amd import field_mm_on_device
    a, b = synth.make_pair("1080p-warp", seed=0)
    p, d, leaf, K = 11, 16, 64, 5
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    comp, mu = _basis(pa[::11], pb[::11], d)
    depth = O.tree_depth(pb.shape[0], leaf)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=1, precision="bf16", return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf_all = tree.natural_leaves(res["red_a"])
    pick = np.random.default_rng(5).choice(1 << depth, 400, replace=False)
    rows = np.nonzero(np.isin(qleaf_all, pick))[0]
    wi, wd = O.leaf_knn_exact(pa[rows], pb, qleaf_all[rows], tree.leaves, tree.orig, 1, K)
    _check_bf16(res, rows, wi, wd)


@pytest.mark.parametrize("p,K,shortlist,nprobe", [(7, 5, 20, 1), (11, 5, 20, 2), (9, 8, 32, 1)])
def test_fp8_shortlist_rerank(p, K, shortlist, nprobe, gpu_available):
    from annfmp_amd import synth
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=30 + p)
    res, pa, pb, tree, rows, qleaf, wi, wd = _run(a, b, p, 8, 32, K, nprobe, "fp8", shortlist)
    knn = res["knn"].astype(np.int64)
    # every entry is a point of a probed leaf, listed once
    leaf_of_b = np.empty(pb.shape[0], np.int64)
    lv = tree.leaves.reshape(-1, 2)
    for l, (f, t) in enumerate(lv):
        leaf_of_b[tree.orig[f:t]] = l
    assert (knn >= 0).all()
    assert ((leaf_of_b[knn] ^ qleaf[:, None]) < nprobe).all()
    assert all(len(set(r)) == K for r in knn[:200])
    # ordered by exact (distance, index)
    d2 = O.exact_pair_d2(pa, pb, np.repeat(rows, K), knn.ravel()).reshape(-1, K)
    key = d2 * (1 << 32) + knn
    assert (np.diff(key, axis=1) > 0).all()
    assert np.array_equal(res["nn"], knn[:, 0].astype(np.int32))
    expect = np.sqrt(d2[:, 0].astype(np.float64)).astype(np.float32)
    assert np.array_equal(res["dist"].view(np.uint32), expect.view(np.uint32))
    # never better than exact, and the exact nearest is found on most queries
    assert (d2 >= wd).all()
    hit = (knn[:, 0] == wi[:, 0]).mean()
    print(f"fp8 p={p} S={shortlist}: exact nearest found on {hit:.4f} of queries")
    assert hit >= 0.8


def test_variant_arguments_rejected(gpu_available):
    from annfmp_amd import field_mm_on_device, synth
    from annfmp_amd._lib import ANNFError
    a, b = synth.make_pair(synth.small_spec(48, 64), seed=1)
    comp = np.eye(4, 75)
    mu = np.zeros(75)
    with pytest.raises(ANNFError):
        field_mm_on_device(a, b, 5, comp, mu, 4, 5, nprobe=3)          # not a power of two
    with pytest.raises(ANNFError):
        field_mm_on_device(a, b, 5, comp, mu, 4, 5, precision="fp8", shortlist=4)   # shortlist < K
    with pytest.raises(ANNFError):
        field_mm_on_device(a, b, 5, comp, mu, 4, 5, nprobe=32)         # more leaves than the tree has


@pytest.mark.parametrize("p,K,seed", [(7, 5, 41), (5, 3, 42), (9, 8, 43)])
def test_bf16_propagation_matches_oracle(p, K, seed, gpu_available):
    """The propagated pass: the K best exact keys over the natural leaf and
    the leaves named by the row above's first-pass matches."""
    from annfmp_amd import synth, field_mm_on_device
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=seed)
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    comp, mu = _basis(pa, pb, 8)
    depth = O.tree_depth(pb.shape[0], 32)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=1, precision="bf16", propagate=True,
                             return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    n_cols = a.shape[1] - p + 1
    wi, wd = O.leaf_knn_propagated(pa, pb, qleaf, tree.leaves, tree.orig, tree.pli, n_cols, K)
    rows = np.arange(pa.shape[0])
    _check_bf16(res, rows, wi, wd)
    # propagation only adds candidates: never worse than the first pass
    i1, d1 = O.leaf_knn_exact(pa, pb, qleaf, tree.leaves, tree.orig, 1, K)
    assert (wd[:, 0] <= d1[:, 0]).all() and (wd[:, 0] < d1[:, 0]).any()


@pytest.mark.parametrize("nprobe,prop", [(2, True), (4, True), (2, 2)])
def test_bf16_propagation_with_several_probed_leaves(nprobe, prop, gpu_available):
    """nprobe > 1 with the propagated pass: a propagated leaf that the first
    pass already scanned (natural leaf ^ r, r < nprobe) is left out, so every
    list holds K distinct B indices and equals the oracle's."""
    from annfmp_amd import synth, field_mm_on_device
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=60 + nprobe)
    p, K = 7, 5
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    comp, mu = _basis(pa, pb, 8)
    depth = O.tree_depth(pb.shape[0], 32)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=nprobe, precision="bf16", propagate=prop,
                             return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    n_cols = a.shape[1] - p + 1
    prop_k = None if prop is True else int(prop)
    wi, wd = O.leaf_knn_propagated(pa, pb, qleaf, tree.leaves, tree.orig, tree.pli, n_cols, K, nprobe=nprobe,
                                   prop_k=prop_k)
    rows = np.arange(pa.shape[0])
    _check_bf16(res, rows, wi, wd)
    knn = res["knn"]
    assert all(len(set(r)) == K for r in knn.tolist())


@pytest.mark.parametrize("nprobe,seed", [(1, 71), (2, 72)])
def test_bf16_row_chain_matches_oracle(nprobe, seed, gpu_available):
    """propagate="chain": rows in order, each row's propagated leaves from the
    row above's final lists (the reference's dependence, base.c:227-306);
    equal to the oracle's chain bit for bit, and it reaches matches the
    parallel pass does not."""
    from annfmp_amd import synth, field_mm_on_device
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=seed)
    p, K = 7, 5
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    comp, mu = _basis(pa, pb, 8)
    depth = O.tree_depth(pb.shape[0], 32)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=nprobe, precision="bf16", propagate="chain",
                             return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    n_cols = a.shape[1] - p + 1
    wi, wd = O.leaf_knn_propagated(pa, pb, qleaf, tree.leaves, tree.orig, tree.pli, n_cols, K, nprobe=nprobe,
                                   chain=True)
    _check_bf16(res, np.arange(pa.shape[0]), wi, wd)
    pi, pd = O.leaf_knn_propagated(pa, pb, qleaf, tree.leaves, tree.orig, tree.pli, n_cols, K, nprobe=nprobe)
    assert (wd[:, 0] < pd[:, 0]).any()


def test_fp8_propagation_properties(gpu_available):
    from annfmp_amd import synth, field_mm_on_device
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=44)
    p, K = 9, 8
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    comp, mu = _basis(pa, pb, 8)
    depth = O.tree_depth(pb.shape[0], 32)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=1, precision="fp8", shortlist=32,
                             propagate=True, return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    n_cols = a.shape[1] - p + 1
    wi, wd = O.leaf_knn_propagated(pa, pb, qleaf, tree.leaves, tree.orig, tree.pli, n_cols, K)
    knn = res["knn"].astype(np.int64)
    rows = np.arange(pa.shape[0])
    d2 = O.exact_pair_d2(pa, pb, np.repeat(rows, K), knn.ravel()).reshape(-1, K)
    assert (np.diff(d2 * (1 << 32) + knn, axis=1) > 0).all()
    assert (d2 >= wd).all()
    hit = (knn[:, 0] == wi[:, 0]).mean()
    print(f"fp8 + propagation: exact propagated nearest found on {hit:.4f} of queries")
    assert hit >= 0.8


def test_front_runs_the_learned_bf16_form(gpu_available):
    """ANNField(leaf_distance="bf16", encoder=...) is the same call as the
    device entry with the encoder's projection."""
    from annfmp_amd import ANNField, synth, field_mm_on_device
    from annfmp_amd.encoder import PatchEncoder
    a, b = synth.make_pair(synth.small_spec(80, 100), seed=50)
    p, K, leaf = 7, 4, 16
    enc = PatchEncoder(p, 16, seed=3)
    f = ANNField(n_neighbors=K, psize=p, dim_reduced=16, algorithm="hip", leaf_size=leaf, propagate=True,
                 verbose=0, leaf_distance="bf16", encoder=enc)
    nn = f.fit(a, b)
    depth = O.tree_depth((b.shape[0] - p + 1) * (b.shape[1] - p + 1), leaf)
    res = field_mm_on_device(a, b, p, None, None, depth, K, precision="bf16", projection=enc, propagate=True)
    assert np.array_equal(nn, res["nn"])
    assert f._model.nn_distances.dtype == np.float32
