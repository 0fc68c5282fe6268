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
    comp, mean = fit_pca_gpu(a, ia, p, d)
    sk = PCA(n_components=d, svd_solver="covariance_eigh").fit(pa[ia].astype(np.float64))
    assert np.array_equal(mean, sk.mean_)
    assert np.array_equal(comp, sk.components_)
