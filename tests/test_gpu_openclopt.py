"""GPU tests of the openclopt-semantics mode (annf_field_openclopt_dev,
annfmp/openclopt/src/futhark/*.fut) against the oracle's restatement
(oracle/ocl_oracle.c) on the same float32 reduced patches: the tree arrays,
the 8-entry lists, their squared distances and the best-of-8 field, bit for
bit."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pair(h, w, p, d, seed, spec=None):
    from annfmp_amd import synth
    from annfmp_amd.openclopt import project_on_device
    from annfmp_amd.encoder import fit_pca
    a, b = synth.make_pair(spec or synth.small_spec(h, w), seed=seed)
    comp, mu = fit_pca(a, p, d, n_samples=4000, seed=seed)
    red_a = project_on_device(a, p, comp, mu)
    red_b = project_on_device(b, p, comp, mu)
    return a, b, red_a, red_b


@pytest.mark.parametrize("m,d,leaf,seed", [(1000, 3, 16, 0), (5, 2, 2, 1), (9, 2, 1, 2), (4001, 6, 32, 3),
                                           (777, 16, 64, 4), (300, 40, 8, 5)])
def test_tree_matches_oracle(m, d, leaf, seed, gpu_available):
    """Padded power-of-two trees (pads, infinite medians, ties) equal the
    oracle's: split dimensions, medians, same-dimension ancestors, order."""
    from annfmp_amd.openclopt import field_openclopt_on_device
    rng = np.random.default_rng(seed)
    X = rng.integers(0, 7, size=(m, d)).astype(np.float32) * np.float32(0.5) - np.float32(1.0)
    X[rng.random(X.shape) < 0.1] = -0.0                     # signed zeros sort before +0
    t = O.OclTree(X, leaf)
    res = field_openclopt_on_device(X[:1], 1, 1, X, leaf, return_tree=True)
    assert res["shape"] == (t.height, t.q, t.ppl, t.mp)
    assert np.array_equal(res["dims"], t.dims)
    assert np.array_equal(res["vals"].view(np.uint32), t.vals.view(np.uint32))
    assert np.array_equal(res["clanc"], t.clanc)
    assert np.array_equal(res["indir"], t.indir)
    assert np.array_equal(res["orig2leaf"], t.orig2leaf)


@pytest.mark.parametrize("h,w,p,d,leaf,seed", [(40, 48, 5, 6, 8, 0), (64, 80, 7, 10, 32, 1), (72, 96, 5, 16, 64, 2)])
def test_field_matches_oracle(h, w, p, d, leaf, seed, gpu_available):
    from annfmp_amd.openclopt import field_openclopt_on_device
    a, b, red_a, red_b = _pair(h, w, p, d, seed)
    nr, nc = h - p + 1, w - p + 1
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    res = field_openclopt_on_device(red_a, nr, nc, red_b, leaf, a, b, p, return_tree=True)
    ref = O.ocl_field(red_a, nr, nc, red_b, leaf, pa, pb)
    assert np.array_equal(res["indir"], ref["indir"])
    assert np.array_equal(res["knn"], ref["knn"])
    assert np.array_equal(res["knd"].view(np.uint32), ref["knd"].view(np.uint32))
    assert np.array_equal(res["nn"], ref["nn"])


def test_field_256_workload_matches_oracle(gpu_available):
    """The 256x256 workload's pair, p = 7, PCA-10, leaf 32 (1024 leaves)."""
    from annfmp_amd import synth
    from annfmp_amd.openclopt import field_openclopt_on_device, project_on_device
    from annfmp_amd.encoder import fit_pca
    a, b = synth.make_pair("256x256-shift", seed=0)
    p, d, leaf = 7, 10, 32
    comp, mu = fit_pca(a, p, d, n_samples=10000, seed=0)
    red_a, red_b = project_on_device(a, p, comp, mu), project_on_device(b, p, comp, mu)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    res = field_openclopt_on_device(red_a, nr, nc, red_b, leaf, a, b, p)
    ref = O.ocl_field(red_a, nr, nc, red_b, leaf, O.extract_patches(a, p), O.extract_patches(b, p))
    assert np.array_equal(res["knn"], ref["knn"])
    assert np.array_equal(res["nn"], ref["nn"])


def test_field_1080p_first_rows_match_oracle(gpu_available):
    """The 1920x1080 pair, p = 11, PCA-16, leaf 64 (32768 leaves of 63,
    m' = 2,064,384): the whole tree and the first 24 rows of the chain (a row
    depends only on the rows above it) equal the oracle's."""
    from annfmp_amd import synth
    from annfmp_amd.openclopt import field_openclopt_on_device, project_on_device
    from annfmp_amd.encoder import fit_pca
    a, b = synth.make_pair("1080p-warp", seed=0)
    p, d, leaf, rows = 11, 16, 64, 24
    comp, mu = fit_pca(a, p, d, seed=0)
    red_a, red_b = project_on_device(a, p, comp, mu), project_on_device(b, p, comp, mu)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    res = field_openclopt_on_device(red_a, nr, nc, red_b, leaf, a, b, p, return_tree=True)
    ref = O.ocl_field(red_a, nr, nc, red_b, leaf, O.extract_patches(a, p), O.extract_patches(b, p), rows=rows)
    assert res["shape"] == ref["shape"] == (14, 32767, 63, 2064384)
    assert np.array_equal(res["dims"], ref["dims"])
    assert np.array_equal(res["vals"].view(np.uint32), ref["vals"].view(np.uint32))
    assert np.array_equal(res["indir"], ref["indir"])
    n = rows * nc
    assert np.array_equal(res["knn"][:n], ref["knn"])
    assert np.array_equal(res["knd"][:n].view(np.uint32), ref["knd"].view(np.uint32))
    assert np.array_equal(res["nn"][:n], ref["nn"])
    assert (res["knn"] >= 0).all()


def test_front_openclopt_semantics(gpu_available):
    """ANNField(semantics="openclopt") = the pipeline on the front's own PCA
    basis; the reference asserts select_best_nn and fixes K = 8."""
    from annfmp_amd import synth
    from annfmp_amd.field import ANNField
    from annfmp_amd.openclopt import field_openclopt
    a, b = synth.make_pair(synth.small_spec(48, 64), seed=6)
    f = ANNField(n_neighbors=3, psize=5, dim_reduced=6, n_subset=2000, leaf_size=16, verbose=0, seed=1,
                 semantics="openclopt")
    nn = f.fit(a, b)
    m = f._model._pca_model
    want = field_openclopt(a, b, 5, m.components_, m.mean_, 16)
    assert np.array_equal(nn, want["nn"])
    assert f._model.knn_indices.shape == (nn.size, 8)
    with pytest.raises(AssertionError):
        ANNField(psize=5, dim_reduced=6, leaf_size=16, verbose=0, select_best_nn=False,
                 semantics="openclopt").fit(a, b)


def test_openclopt_argument_errors(gpu_available):
    from annfmp_amd import _lib
    from annfmp_amd.openclopt import field_openclopt_on_device
    X = np.random.default_rng(0).normal(size=(100, 4)).astype(np.float32)
    with pytest.raises(_lib.ANNFError, match="EUNSUPPORTED"):
        field_openclopt_on_device(X[:50], 5, 10, X, 8)          # row chain with n_b != n_a
    with pytest.raises(_lib.ANNFError, match="EINVAL"):
        field_openclopt_on_device(X, 10, 10, X, 0)              # leaf size 0
