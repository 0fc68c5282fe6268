"""Host-side logic and the C ABI surface (no GPU needed)."""
import math
import os
import re
import warnings

import numpy as np
import pytest

from conftest import ROOT, golden
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    from annfmp_amd import _lib
    header = open(os.path.join(ROOT, "include", "annf_hip.h")).read()
    declared = sorted(set(re.findall(r"\b(annf_[a-z0-9_]+)\s*\(", header)))
    assert set(declared) == set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_library_is_gfx950_code_object():
    from annfmp_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_host_norm_array_matches_reference():
    from annfmp_amd import _lib
    g = golden("norm_array.npz")
    L = _lib.lib()
    out = np.array([L.annf_norm_array(_lib.ptr(np.ascontiguousarray(a), _lib._u8p),
                                      _lib.ptr(np.ascontiguousarray(b), _lib._u8p), a.size)
                    for a, b in zip(g["A"], g["B"])], np.float32)
    assert np.array_equal(out.view(np.uint32), g["out"].view(np.uint32))


@pytest.mark.parametrize("n,leaf", [(2, 2), (100, 3), (62500, 32), (2043700, 64), (8246464, 128), (5, 64)])
def test_tree_depth_rule_matches_oracle(n, leaf):
    from annfmp_amd.kdtree import tree_depth_for
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        assert tree_depth_for(n, leaf) == O.tree_depth(n, leaf)


def test_tree_depth_errors_and_explicit_depth():
    from annfmp_amd.kdtree import KDTree
    t = KDTree(leaf_size=1)
    with pytest.raises(Exception, match="leaf_size"):
        t._compute_final_tree_depth(100, 1, None)
    with pytest.warns(UserWarning):
        assert t._compute_final_tree_depth(100, 1, 50) == 6
    assert t._compute_final_tree_depth(100, 1, 3) == 3


def test_kdtree_params_and_float_type():
    from annfmp_amd.kdtree import KDTree
    t = KDTree()
    assert t.get_params() == {"n_neighbors": 5, "leaf_size": 30, "max_leaves": None, "tree_depth": None,
                              "float_type": "double", "splitting_type": "cyclic", "n_jobs": 1, "verbose": 0}
    with pytest.raises(Exception, match="float32"):
        KDTree(float_type="double").fit(np.zeros((10, 2), np.float32))


def test_annfield_params_and_algorithms():
    from annfmp_amd import ANNField
    f = ANNField()
    p = f.get_params()
    assert p["algorithm"] == "hip" and p["psize"] == 8 and p["dim_reduced"] == 24
    f.set_params(psize=5, n_neighbors=3)
    assert f.psize == 5 and f.n_neighbors == 3
    img = np.zeros((20, 20, 3), np.uint8)
    with pytest.raises(Exception, match="Unknown algorithm"):
        ANNField(algorithm="bogus").fit(img, img)
    with pytest.raises(Exception, match="reference package"):
        ANNField(algorithm="openmp").fit(img, img)


def test_backend_checks_leaf_size_against_k():
    from annfmp_amd import ANNFieldPropKDTreeHIP
    img = np.random.default_rng(0).integers(0, 256, (24, 24, 3)).astype(np.uint8)
    m = ANNFieldPropKDTreeHIP(n_neighbors=8, leaf_size=8, psize=3, dim_reduced=2, n_subset=100, verbose=0)
    with pytest.raises(AssertionError):
        m.fit(img, img)
    with pytest.raises(Exception, match="double"):
        ANNFieldPropKDTreeHIP(float_type="double", verbose=0).fit(img, img)


def test_synthetic_pairs_are_deterministic():
    from annfmp_amd import synth
    a1, b1 = synth.make_pair("256x256-shift", seed=3)
    a2, b2 = synth.make_pair("256x256-shift", seed=3)
    a3, _ = synth.make_pair("256x256-shift", seed=4)
    assert a1.dtype == np.uint8 and a1.shape == (256, 256, 3)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and not np.array_equal(a1, a3)


def test_shift_pair_closed_form():
    """B is A rolled by (+17, -9) plus N(0, 4) noise."""
    from annfmp_amd import synth
    a, b = synth.make_pair("256x256-shift", seed=0)
    diff = b.astype(np.int32) - np.roll(a, (17, -9), axis=(0, 1)).astype(np.int32)
    assert abs(diff.mean()) < 0.2 and 3.0 < diff.std() < 4.5


def test_warp_pair_block_offsets():
    """Away from clipping, each warp block of B is a translated copy of A."""
    from annfmp_amd import synth
    spec = synth.small_spec(90, 120, grid=(4, 3), max_offset=6)
    rng = np.random.default_rng(0)
    a = synth.make_image_a(spec, rng)
    b, offs = synth._warp(a, spec, rng)
    dy, dx = offs[1, 2]
    assert np.array_equal(b[30:60, 60:90], a[(np.arange(30, 60) + dy) % 90][:, (np.arange(60, 90) + dx) % 120])


def test_leaf_range_formula_matches_oracle_tree():
    """The kernels compute a leaf's [from, to) from (n, depth) alone: each
    segment splits at floor(c/2).  Check that formula against built trees."""
    def leaf_range(l, n, depth):
        lo, hi = 0, n
        for lev in range(depth - 1, -1, -1):
            mid = lo + (hi - lo) // 2
            if (l >> lev) & 1:
                lo = mid
            else:
                hi = mid
        return lo, hi
    rng = np.random.default_rng(0)
    for n, leaf in [(777, 5), (100003, 64), (62500, 32), (4099, 2)]:
        depth = O.tree_depth(n, leaf)
        t = O.OracleTree(rng.standard_normal((n, 2)).astype(np.float32), depth)
        for l in range(1 << depth):
            assert leaf_range(l, n, depth) == (t.leaves[2 * l], t.leaves[2 * l + 1])


def test_sklearn_sample_indices_match_extract_patches_2d():
    """The GPU PCA fit samples the positions extract_patches_2d(max_patches=n,
    random_state=seed) samples (openmp/base.py:170-177)."""
    import numpy as np
    from sklearn.feature_extraction import image
    from annfmp_amd.pca import sklearn_sample_indices
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    p = 7
    nc = 53 - p + 1
    for n, seed in ((500, 0), (123, 9), (5000, 3)):
        want = image.extract_patches_2d(img, (p, p), max_patches=n, random_state=seed).reshape(-1, p * p * 3)
        idx = sklearn_sample_indices(37, 53, p, n, seed)
        got = np.stack([img[i // nc:i // nc + p, i % nc:i % nc + p].reshape(-1) for i in idx])
        assert np.array_equal(got, want)
