"""GPU tests of the workloads at their full sizes: the trained encoder on the
1920x1080 pairs (and what it gains over PCA), FP8 leaf distances at 1080p,
the 3840x2160 combined form, and the image-level front against the oracle
on the same images.

Exactness bars: BF16 full-patch lists are exact integer k-NN, so they must
equal the oracle's (leaf_knn_exact / leaf_knn_propagated_rows) bit for bit;
FP8 lists are approximate and are checked by properties (members of the
scanned leaves, exactly ordered by (exact distance, index), never better
than the exact answer over the same leaves, exact nearest found on most
queries).
"""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ENC11 = os.path.join(ROOT, "annfmp_amd", "data", "encoder_p11_d16.npz")
ENC9 = os.path.join(ROOT, "annfmp_amd", "data", "encoder_p9_d16.npz")


def _leaf_of_points(tree, n_b):
    leaf_of_b = np.empty(n_b, np.int64)
    lv = tree.leaves.reshape(-1, 2)
    for l, (f, t) in enumerate(lv):
        leaf_of_b[tree.orig[f:t]] = l
    return leaf_of_b


def _check_exact_lists(res, rows, wi, wd):
    assert np.array_equal(res["knn"][rows], wi.astype(np.int32))
    assert np.array_equal(res["nn"][rows], wi[:, 0].astype(np.int32))
    expect = np.sqrt(wd[:, 0].astype(np.float64)).astype(np.float32)
    assert np.array_equal(res["dist"][rows].view(np.uint32), expect.view(np.uint32))


def _check_approx_lists(res, pa, pb, rows, K, exact_d):
    """FP8 lists: K distinct entries ordered by exact (distance, index), the
    distance field their exact root, never better than the exact answer;
    returns the fraction of rows whose first entry is the exact nearest."""
    knn = res["knn"][rows].astype(np.int64)
    assert (knn >= 0).all()
    d2 = O.exact_pair_d2(pa, pb, np.repeat(rows, K), knn.ravel()).reshape(-1, K)
    assert (np.diff(d2 * (1 << 32) + knn, axis=1) > 0).all()
    assert np.array_equal(res["nn"][rows], knn[:, 0].astype(np.int32))
    expect = np.sqrt(d2[:, 0].astype(np.float64)).astype(np.float32)
    assert np.array_equal(res["dist"][rows].view(np.uint32), expect.view(np.uint32))
    assert (d2 >= exact_d).all()
    return float((d2[:, 0] == exact_d[:, 0]).mean())


# ---------------------------------------------------------------------------
# the learned encoder on the 1920x1080 pairs
# ---------------------------------------------------------------------------

def test_trained_encoder_1080p_bf16_matches_exact_leaf_knn(gpu_available):
    """The committed p = 11 encoder (trained by the library's own kernels) as
    the tree projection of the 1920x1080 pair, seed 0, leaf 64, K = 5: the
    BF16 lists of the queries of 400 sampled leaves equal the exact k-NN over
    their natural leaf of the tree built on the encoder's embeddings (the GPU
    tree is rebuilt by the oracle from the same embeddings)."""
    from annfmp_amd import synth, field_mm_on_device
    from annfmp_amd.encoder import PatchEncoder
    enc = PatchEncoder.load(ENC11)
    assert (enc.psize, enc.d, enc.h1, enc.h2) == (11, 16, 128, 128)
    a, b = synth.make_pair("1080p-warp", seed=0)
    p, K = 11, 5
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    depth = O.tree_depth(pb.shape[0], 64)
    assert depth == 14
    res = field_mm_on_device(a, b, p, None, None, depth, K, nprobe=1, precision="bf16", projection=enc,
                             return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    pick = np.random.default_rng(3).choice(1 << depth, 400, replace=False)
    rows = np.nonzero(np.isin(qleaf, pick))[0]
    assert rows.size > 20000
    wi, wd = O.leaf_knn_exact(pa[rows], pb, qleaf[rows], tree.leaves, tree.orig, 1, K)
    _check_exact_lists(res, rows, wi, wd)


@pytest.mark.parametrize("seed", [1, 2])
def test_trained_encoder_gains_over_pca_on_held_out_pairs(seed, gpu_available):
    """What the trained network gains over the method without it: on held-out
    1920x1080 pairs (training used the 512x512 pairs of seeds 1000-1063),
    recall@5 of the BF16 leaf k-NN against the exact field on 8192 sampled
    queries, with the encoder's tree against PCA-16's at equal leaf visits
    (one leaf of 124-125 points per query).  The measured gain is several
    points of recall (0.13-0.16 against 0.12-0.14); the test asserts the
    encoder is ahead on both recall and mean distance ratio."""
    from annfmp_amd.encoder import PatchEncoder, compare_with_pca
    enc = PatchEncoder.load(ENC11)
    row = compare_with_pca(enc, [seed])[0]
    print("encoder vs PCA-16:", row)
    assert row["encoder"]["recall@5"] > row["pca"]["recall@5"]
    assert row["encoder"]["distance_ratio"] < row["pca"]["distance_ratio"]


def test_front_falls_back_to_pca_when_the_encoder_loses(gpu_available):
    """The front's PCA fallback: with encoder_check > 0 it measures recall@k
    of the encoder's field and of PCA's on sampled queries against the exact
    field and keeps the better one.  An untrained (random) encoder loses to
    PCA by a wide margin (measured 0.48 against 0.73), so the returned field is
    PCA's field, entry for entry."""
    from annfmp_amd import ANNField, synth
    from annfmp_amd.encoder import PatchEncoder
    a, b = synth.make_pair(synth.small_spec(200, 260, n_rects=10), seed=70)
    p, K = 11, 5
    rand = PatchEncoder(p, 16, seed=4)
    f = ANNField(n_neighbors=K, psize=p, dim_reduced=16, n_subset=4000, algorithm="hip", leaf_size=64,
                 verbose=0, leaf_distance="bf16", encoder=rand, encoder_check=2048)
    nn = f.fit(a, b)
    m = f._model
    print("fallback check:", m.encoder_check_result)
    assert m.projection_used in ("pca", "encoder")
    r = m.encoder_check_result
    assert m.projection_used == ("encoder" if r["encoder"] >= r["pca"] else "pca")
    assert r["pca"] > r["encoder"] + 0.1 and m.projection_used == "pca"
    # the returned field is the PCA front's own
    f_pca = ANNField(n_neighbors=K, psize=p, dim_reduced=16, n_subset=4000, algorithm="hip", leaf_size=64,
                     verbose=0, leaf_distance="bf16")
    assert np.array_equal(nn, f_pca.fit(a, b))


# ---------------------------------------------------------------------------
# FP8 leaf distances at 1080p (the precision variant of the 1080p workload)
# ---------------------------------------------------------------------------

def test_fp8_1080p_shortlist20(gpu_available):
    """1920x1080 pair, PCA-16, leaf 64, K = 5, FP8-E4M3 operands with
    per-patch scales and a top-20 exact re-rank; queries of 300 sampled
    leaves against the exact k-NN over their leaf."""
    from annfmp_amd import synth, field_mm_on_device
    from annfmp_amd.encoder import fit_pca
    a, b = synth.make_pair("1080p-warp", seed=0)
    p, d, K = 11, 16, 5
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    comp, mu = fit_pca(a, p, d, seed=0)
    depth = O.tree_depth(pb.shape[0], 64)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=1, precision="fp8", shortlist=20,
                             return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    pick = np.random.default_rng(4).choice(1 << depth, 300, replace=False)
    rows = np.nonzero(np.isin(qleaf, pick))[0]
    wi, wd = O.leaf_knn_exact(pa[rows], pb, qleaf[rows], tree.leaves, tree.orig, 1, K)
    leaf_of_b = _leaf_of_points(tree, pb.shape[0])
    assert (leaf_of_b[res["knn"][rows]] == qleaf[rows][:, None]).all()
    hit = _check_approx_lists(res, pa, pb, rows, K, wd)
    print(f"fp8 1080p top-20: exact nearest found on {hit:.4f} of {rows.size} queries")
    assert hit >= 0.9


# ---------------------------------------------------------------------------
# the 3840x2160 combined form
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def pair_4k():
    from annfmp_amd import synth
    a, b = synth.make_pair("4k-warp", seed=0)
    p = 9
    return a, b, O.extract_patches(a, p), O.extract_patches(b, p)


def _rows_4k(n_rows, n_cols, n):
    """n A patches spread over the image, with whole runs down columns (so
    a propagated chain is checked row after row)."""
    rng = np.random.default_rng(17)
    cols = rng.choice(n_cols, n // 40, replace=False)
    r0 = rng.integers(1, n_rows - 40, cols.size)
    rows = (r0[:, None] + np.arange(40)[None, :]) * n_cols + cols[:, None]
    return np.unique(rows.ravel())


def test_4k_learned_bf16_propagation_matches_oracle(pair_4k, gpu_available):
    """3840x2160, p = 9 learned encoder (243-128-128-16), leaf 128 (depth
    15), K = 8, BF16 leaf k-NN with the propagated pass: 4000 sampled queries
    equal the oracle's exact k-NN over their natural leaf and the leaves the
    row above's first-pass matches propagate."""
    from annfmp_amd import field_mm_on_device
    from annfmp_amd.encoder import PatchEncoder
    a, b, pa, pb = pair_4k
    enc = PatchEncoder.load(ENC9)
    p, K = 9, 8
    depth = O.tree_depth(pb.shape[0], 128)
    assert depth == 15
    res = field_mm_on_device(a, b, p, None, None, depth, K, nprobe=1, precision="bf16", projection=enc,
                             propagate=True, return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    n_cols = a.shape[1] - p + 1
    rows = _rows_4k(a.shape[0] - p + 1, n_cols, 4000)
    wi, wd = O.leaf_knn_propagated_rows(pa, pb, qleaf, tree.leaves, tree.orig, tree.pli, n_cols, K, rows)
    _check_exact_lists(res, rows, wi, wd)


def test_4k_learned_fp8_top32_propagation_properties(pair_4k, gpu_available):
    """The combined form itself: FP8-E4M3 leaf distances with per-patch
    scales, top-32 exact re-rank and the propagated pass, at 3840x2160 with
    the p = 9 encoder, leaf 128, K = 8.  Checked by properties on 4000
    sampled queries: the natural-leaf pass's FP8 answer is never better than
    the exact one over the natural leaf plus what propagation adds, lists
    are exactly ordered, and the exact nearest over the natural leaf is
    found or beaten on most queries."""
    from annfmp_amd import field_mm_on_device
    from annfmp_amd.encoder import PatchEncoder
    a, b, pa, pb = pair_4k
    enc = PatchEncoder.load(ENC9)
    p, K = 9, 8
    depth = O.tree_depth(pb.shape[0], 128)
    res = field_mm_on_device(a, b, p, None, None, depth, K, nprobe=1, precision="fp8", shortlist=32,
                             projection=enc, propagate=True, return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    n_cols = a.shape[1] - p + 1
    rows = _rows_4k(a.shape[0] - p + 1, n_cols, 4000)
    knn = res["knn"][rows].astype(np.int64)
    d2 = O.exact_pair_d2(pa, pb, np.repeat(rows, K), knn.ravel()).reshape(-1, K)
    assert (knn >= 0).all()
    assert (np.diff(d2 * (1 << 32) + knn, axis=1) > 0).all()
    expect = np.sqrt(d2[:, 0].astype(np.float64)).astype(np.float32)
    assert np.array_equal(res["dist"][rows].view(np.uint32), expect.view(np.uint32))
    # the natural leaf is always scanned: its exact nearest is found or beaten
    wi, wd = O.leaf_knn_exact(pa[rows], pb, qleaf[rows], tree.leaves, tree.orig, 1, K)
    ok = (d2[:, 0] <= wd[:, 0]).mean()
    print(f"4k fp8 top-32 + propagation: natural-leaf exact nearest found or beaten on {ok:.4f}")
    assert ok >= 0.9


# ---------------------------------------------------------------------------
# the image-level front against the oracle on the same images
# ---------------------------------------------------------------------------

def _front_vs_oracle(a, b, p, d, leaf, K, n_subset, projection="fp64"):
    """ANNField(algorithm="hip", projection=projection) end to end on the two
    images, and the oracle on the same images and the same fitted PCA:
    oracle.project_f64 (scikit-learn's FP64 transform then astype(float32),
    openmp/base.py:193-200), the oracle's tree and rows.  Returns the
    comparison figures."""
    from annfmp_amd import ANNField
    f = ANNField(n_neighbors=K, psize=p, dim_reduced=d, n_subset=n_subset, algorithm="hip", leaf_size=leaf,
                 verbose=0, projection=projection)
    nn = f.fit(a, b)
    pca = f._model._pca_model
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    ra = O.project_f64(pa, pca.components_, pca.mean_)
    rb = O.project_f64(pb, pca.components_, pca.mean_)
    depth = O.tree_depth(pb.shape[0], leaf)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    ref = O.OracleTree(rb, depth)
    final, _, _ = ref.process_rows(ra, nr, nc, K, True, True, pa, pb)
    # the front's own projected arrays, through the device entry with the same basis
    from annfmp_amd import field_on_device
    res = field_on_device(a, b, p, pca.components_, pca.mean_, depth, K, projection=projection, return_reduced=True)
    assert np.array_equal(res["nn"], nn)
    def ulp(x, y):   # |x - y| in units in the last place of the larger magnitude (float32)
        big = np.maximum(np.abs(x), np.abs(y)).astype(np.float32)
        return np.abs(x.astype(np.float64) - y.astype(np.float64)) / np.spacing(big).astype(np.float64)
    ua, ub = ulp(res["red_a"], ra), ulp(res["red_b"], rb)
    # a difference of the two FP64 summation orders is about 1e-16 of the
    # largest terms, i.e. about 1e-8 of the largest projected value: one ulp of
    # a large value, a few ulps of a value near zero
    scale = float(max(np.abs(ra).max(), np.abs(rb).max()))
    absd = max(np.abs(res["red_a"].astype(np.float64) - ra).max(), np.abs(res["red_b"].astype(np.float64) - rb).max())
    if projection == "fp64":
        assert absd <= 1e-7 * scale, (absd, scale)
    mism = np.nonzero(nn != final)[0]
    out = {"n": nn.size, "nn_mismatch": int(mism.size), "red_ulp_max": float(max(ua.max(), ub.max())),
           "red_differ": int((ua > 0).sum() + (ub > 0).sum()), "red_values": int(ua.size + ub.size),
           "abs_diff_max_over_scale": float(absd / scale)}
    if mism.size:
        # the two chosen matches' full-patch distances (norm_array)
        dg = np.array([O.norm_array(pa[i], pb[nn[i]]) for i in mism], np.float64)
        do = np.array([O.norm_array(pa[i], pb[final[i]]) for i in mism], np.float64)
        out["rel_gap_max"] = float(np.max(np.abs(dg - do) / np.maximum(do, 1e-30)))
    if out["red_differ"] == 0:
        # identical reduced arrays: identical tree and field
        assert mism.size == 0
    print("front (%s) vs oracle:" % projection, out)
    return out, res, ra, rb


def test_front_fp64_matches_oracle_256(gpu_available):
    """The 256x256 shift pair (p = 7, PCA-10 from 10,000 sampled patches,
    leaf 32, K = 1, propagation and best-of-K on): the HIP front's FP64
    projection is the reference's transform up to the last bit of a few
    values, and the NNF equals the oracle's on the same images."""
    from annfmp_amd import synth
    a, b = synth.make_pair("256x256-shift", seed=0)
    out, res, ra, rb = _front_vs_oracle(a, b, 7, 10, 32, 1, 10000)
    assert out["red_differ"] <= 1e-5 * out["red_values"]
    assert out["nn_mismatch"] <= 1e-4 * out["n"]


def test_front_fp64_matches_oracle_1080p(gpu_available):
    """The 1920x1080 pair (p = 11, PCA-16 from 50,000 sampled patches, leaf
    64, K = 5): the FP64 projection differs from the reference's transform in
    about 1e-6 of the values (measured 102 of 65.4 M, by at most 1e-7 of the
    largest value), and the NNF equals the oracle's on the same images (0
    differing patches measured; the bound allows 1e-4)."""
    from annfmp_amd import synth
    a, b = synth.make_pair("1080p-warp", seed=0)
    out, res, ra, rb = _front_vs_oracle(a, b, 11, 16, 64, 5, 50000)
    assert out["red_differ"] <= 1e-5 * out["red_values"]
    assert out["nn_mismatch"] <= 1e-4 * out["n"]


@pytest.mark.parametrize("workload", ["256", "1080p"])
def test_front_fp32_projection_against_the_reference_transform(workload, gpu_available):
    """The default front projects in FP32 (summation order of the MFMA GEMM),
    the reference in FP64 then casts (openmp/base.py:193-200): how many NNF
    entries that changes on the same images and basis.  A reduced value moves
    by a few FP32 ulps, which reorders near-equal keys of the tree build and
    near-equal candidates, and the row chain carries a changed match down its
    column.  Measured: 256x256 0 of 62,500 patches, 1080p 15,194 of 2,043,700
    (0.74 %), where the two chosen matches' full-patch distances differ by up
    to 0.81 relative (a changed match travels down the column); the reduced
    values differ by at most 1.0e-6 of the largest."""
    from annfmp_amd import synth
    if workload == "256":
        a, b = synth.make_pair("256x256-shift", seed=0)
        out, res, ra, rb = _front_vs_oracle(a, b, 7, 10, 32, 1, 10000, projection="fp32")
    else:
        a, b = synth.make_pair("1080p-warp", seed=0)
        out, res, ra, rb = _front_vs_oracle(a, b, 11, 16, 64, 5, 50000, projection="fp32")
    rate = out["nn_mismatch"] / out["n"]
    print(f"fp32 front, {workload}: {out['nn_mismatch']} of {out['n']} NNF entries differ ({100 * rate:.3f} %)")
    assert out["abs_diff_max_over_scale"] <= 2e-5
    assert rate <= 0.05


def test_front_runs_the_reference_example_configuration(gpu_available):
    """The reference's example call (examples/simple.py:16-28: n_neighbors 8,
    psize 8, dim_reduced 40, leaf_size 128, propagation and best-of-K on)
    through the HIP front on a crop of its example photographs, FP64
    projection: the field equals the oracle's on the same images."""
    from conftest import golden
    g = golden("example_crop.npz")
    out, res, ra, rb = _front_vs_oracle(g["a"], g["b"], 8, 40, 128, 8, 1000)
    assert out["nn_mismatch"] == 0 or out["red_differ"] > 0
    assert out["nn_mismatch"] <= 1e-4 * out["n"]
