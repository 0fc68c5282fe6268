"""PCA of sampled patches with the FP64 covariance accumulated on the GPU.

``pca_covariance`` returns scikit-learn's covariance_eigh matrix (1.7.2,
decomposition/_pca.py: ``C = X.T @ X; C -= (n * mean_col) * mean_row;
C /= n - 1``) of the listed A and B patches, computed by
``annf_pca_cov_dev``: the sums are exact integers, so the matrix equals the
CPU one bit for bit.  ``fit_pca_gpu`` runs the whole fit on the GPU
(``annf_pca_fit_dev``: the covariance, an FP64 eigensolve by rocSOLVER's
dsyevd, the leading components with sklearn's sign rule -- largest |loading|
of each component positive, svd_flip with u_based_decision=False); with
``eigensolve="host"`` it finishes with LAPACK's eigh on the host instead,
which gives sklearn's PCA(svd_solver="covariance_eigh") fit bit for bit.
``sklearn_sample_indices`` reproduces the positions
``extract_patches_2d(max_patches=n, random_state=seed)`` samples
(sklearn 1.7.2 feature_extraction/image.py: ``randint`` of the rows, then
of the columns, from one RandomState).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import device as dev


def pca_covariance(image_a, idx_a, image_b=None, idx_b=None, psize=8, device=0, handle=None):
    """(mean [D], covariance [D, D]) float64 of patches idx_a of image_a
    followed by patches idx_b of image_b (indices in patch order)."""
    torch = dev.torch()
    cuda = dev.cuda_device(device)
    h = handle or _lib.Handle(device)
    p = int(psize)
    D = p * p * 3
    ia = np.ascontiguousarray(idx_a if idx_a is not None else [], dtype=np.int32)
    ib = np.ascontiguousarray(idx_b if idx_b is not None else [], dtype=np.int32)
    ta = torch.from_numpy(np.ascontiguousarray(image_a, dtype=np.uint8)).to(cuda)
    tb = torch.from_numpy(np.ascontiguousarray(image_b if image_b is not None else image_a, dtype=np.uint8)).to(cuda)
    tia = torch.from_numpy(ia).to(cuda) if ia.size else None
    tib = torch.from_numpy(ib).to(cuda) if ib.size else None
    mean = torch.empty(D, dtype=torch.float64, device=cuda)
    cov = torch.empty((D, D), dtype=torch.float64, device=cuda)
    ha, wa = image_a.shape[:2]
    hb, wb = (image_b if image_b is not None else image_a).shape[:2]
    _lib.check("annf_pca_cov_dev", _lib.lib().annf_pca_cov_dev(
        h.h, dev.dptr(ta), ha, wa, dev.dptr(tia), int(ia.size), dev.dptr(tb), hb, wb, dev.dptr(tib), int(ib.size), p,
        dev.dptr(mean), dev.dptr(cov)))
    return mean.cpu().numpy(), cov.cpu().numpy()


def components_from_covariance(cov, d):
    """The d leading eigenvectors (rows) of the covariance, sklearn's order
    and signs."""
    _, V = np.linalg.eigh(cov)
    comp = V[:, ::-1][:, :d].T.copy()
    comp *= np.sign(comp[np.arange(d), np.abs(comp).argmax(1)])[:, None]
    return comp


def sklearn_sample_indices(h, w, p, n, seed):
    """Patch indices (row * n_cols + col) of extract_patches_2d(img, (p, p),
    max_patches=n, random_state=seed) on an h x w image."""
    nr, nc = h - p + 1, w - p + 1
    n = min(int(n), nr * nc)
    rng = np.random.RandomState(seed)
    i_s = rng.randint(nr, size=n)
    j_s = rng.randint(nc, size=n)
    return (i_s.astype(np.int64) * nc + j_s).astype(np.int32)


def fit_pca_gpu(image_a, idx_a, psize, d, image_b=None, idx_b=None, device=0, handle=None, eigensolve="gpu",
                return_variance=False):
    """(components [d, D], mean [D]) of the listed patches (A's, then B's).
    eigensolve="gpu": everything on the GPU (annf_pca_fit_dev); "host": the
    GPU covariance, then LAPACK eigh on the host."""
    if eigensolve == "host":
        mean, cov = pca_covariance(image_a, idx_a, image_b, idx_b, psize, device, handle)
        comp = components_from_covariance(cov, d)
        if return_variance:
            w = np.linalg.eigvalsh(cov)[::-1][:d]
            return comp, mean, np.maximum(w, 0.0)
        return comp, mean
    if eigensolve != "gpu":
        raise ValueError("eigensolve must be 'gpu' or 'host'")
    torch = dev.torch()
    cuda = dev.cuda_device(device)
    h = handle or _lib.Handle(device)
    p = int(psize)
    D = p * p * 3
    ia = np.ascontiguousarray(idx_a if idx_a is not None else [], dtype=np.int32)
    ib = np.ascontiguousarray(idx_b if idx_b is not None else [], dtype=np.int32)
    ta = torch.from_numpy(np.ascontiguousarray(image_a, dtype=np.uint8)).to(cuda)
    tb = torch.from_numpy(np.ascontiguousarray(image_b if image_b is not None else image_a, dtype=np.uint8)).to(cuda)
    tia = torch.from_numpy(ia).to(cuda) if ia.size else None
    tib = torch.from_numpy(ib).to(cuda) if ib.size else None
    comp = torch.empty((d, D), dtype=torch.float64, device=cuda)
    mean = torch.empty(D, dtype=torch.float64, device=cuda)
    var = torch.empty(d, dtype=torch.float64, device=cuda)
    ha, wa = image_a.shape[:2]
    hb, wb = (image_b if image_b is not None else image_a).shape[:2]
    _lib.check("annf_pca_fit_dev", _lib.lib().annf_pca_fit_dev(
        h.h, dev.dptr(ta), ha, wa, dev.dptr(tia), int(ia.size), dev.dptr(tb), hb, wb, dev.dptr(tib), int(ib.size), p,
        int(d), dev.dptr(comp), dev.dptr(mean), dev.dptr(var)))
    if return_variance:
        return comp.cpu().numpy(), mean.cpu().numpy(), var.cpu().numpy()
    return comp.cpu().numpy(), mean.cpu().numpy()
