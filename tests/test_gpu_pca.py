"""GPU tests of the FP64 PCA covariance (annf_pca_cov_dev) against
scikit-learn 1.7.2's covariance_eigh arithmetic (decomposition/_pca.py),
bit for bit, and of the fitted components against sklearn's PCA."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sklearn_cov(X):
    X = X.astype(np.float64)
    n = X.shape[0]
    mean = np.mean(X, axis=0)
    C = X.T @ X
    C -= n * np.reshape(mean, (-1, 1)) * np.reshape(mean, (1, -1))
    C /= n - 1
    return mean, C


@pytest.mark.parametrize("p,na,nb", [(5, 3000, 0), (7, 1500, 1700), (11, 5000, 5000)])
def test_covariance_bit_exact(p, na, nb, gpu_available):
    from annfmp_amd import synth
    from annfmp_amd.pca import pca_covariance
    a, b = synth.make_pair(synth.small_spec(120, 150), seed=p)
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    rng = np.random.default_rng(p)
    ia = rng.integers(0, pa.shape[0], na)          # with replacement, as extract_patches_2d samples
    ib = rng.integers(0, pb.shape[0], nb)
    mean, C = pca_covariance(a, ia, b if nb else None, ib if nb else None, p)
    wm, wC = _sklearn_cov(np.concatenate([pa[ia], pb[ib]]))
    assert np.array_equal(mean.view(np.uint64), wm.view(np.uint64))
    assert np.array_equal(C.view(np.uint64), wC.view(np.uint64))


def test_components_match_sklearn(gpu_available):
    from sklearn.decomposition import PCA
    from annfmp_amd import synth
    from annfmp_amd.pca import fit_pca_gpu
    a, b = synth.make_pair(synth.small_spec(120, 150), seed=3)
    p, d = 7, 10
    pa = O.extract_patches(a, p)
    ia = np.random.default_rng(3).integers(0, pa.shape[0], 10000)
    comp, mean = fit_pca_gpu(a, ia, p, d, eigensolve="host")   # GPU covariance, LAPACK eigh
    sk = PCA(n_components=d, svd_solver="covariance_eigh").fit(pa[ia].astype(np.float64))
    assert np.array_equal(mean, sk.mean_)
    assert np.array_equal(comp, sk.components_)


@pytest.mark.parametrize("p,d,na,nb", [(7, 10, 10000, 0), (11, 16, 25000, 25000), (8, 40, 6000, 6000)])
def test_gpu_eigensolve_matches_sklearn(p, d, na, nb, gpu_available):
    """annf_pca_fit_dev (covariance and rocSOLVER eigensolve on the GPU) gives
    sklearn's PCA(svd_solver="covariance_eigh") fit: the mean bit for bit, the
    components to 1e-9 (the eigensolvers round differently; the components
    are unit vectors) with the same signs, the variances to 1e-10 relative."""
    from sklearn.decomposition import PCA
    from annfmp_amd import synth
    from annfmp_amd.pca import fit_pca_gpu
    a, b = synth.make_pair(synth.small_spec(160, 200), seed=p)
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    rng = np.random.default_rng(p + d)
    ia = rng.integers(0, pa.shape[0], na)
    ib = rng.integers(0, pb.shape[0], nb)
    comp, mean, var = fit_pca_gpu(a, ia, p, d, b if nb else None, ib if nb else None, return_variance=True)
    X = np.concatenate([pa[ia], pb[ib]]).astype(np.float64)
    sk = PCA(n_components=d, svd_solver="covariance_eigh").fit(X)
    assert np.array_equal(mean, sk.mean_)
    # the leading eigenvalues of these patch sets are well separated, so each
    # component is determined up to its sign, which both sides fix the same way
    gaps = np.diff(sk.explained_variance_[::-1])
    assert gaps.min() > 1e-6 * sk.explained_variance_[0]
    assert np.abs(comp - sk.components_).max() < 1e-9
    assert np.allclose(var, sk.explained_variance_, rtol=1e-10, atol=0)


def test_backend_gpu_pca_fit_matches_sklearn_fit(gpu_available):
    """ANNField(pca_fit="gpu") fits the reference's own samples
    (extract_patches_2d with seed and seed + 1, openmp/base.py:168-191) on the
    GPU: its basis equals sklearn's covariance_eigh fit of those samples."""
    from sklearn.decomposition import PCA
    from sklearn.feature_extraction import image
    from annfmp_amd import synth
    from annfmp_amd.field import ANNField
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=2)
    p, d, n_subset, seed = 5, 8, 4000, 3
    f = ANNField(n_neighbors=2, psize=p, dim_reduced=d, n_subset=n_subset, leaf_size=16, verbose=0, seed=seed,
                 pca_fit="gpu")
    nn = f.fit(a, b)
    sa = image.extract_patches_2d(a, (p, p), max_patches=n_subset // 2, random_state=seed).reshape(n_subset // 2, -1)
    sb = image.extract_patches_2d(b, (p, p), max_patches=n_subset // 2, random_state=seed + 1).reshape(n_subset // 2, -1)
    sk = PCA(n_components=d, svd_solver="covariance_eigh").fit(np.concatenate([sa, sb]).astype(np.float64))
    m = f._model._pca_model
    assert np.array_equal(m.mean_, sk.mean_)
    assert np.abs(m.components_ - sk.components_).max() < 1e-9
    assert nn.shape == ((96 - p + 1) * (128 - p + 1),) and nn.dtype == np.int32
