"""PCA of sampled patches with the FP64 covariance accumulated on the GPU.

``pca_covariance`` returns scikit-learn's covariance_eigh matrix (1.7.2,
decomposition/_pca.py: ``C = X.T @ X; C -= (n * mean_col) * mean_row;
C /= n - 1``) of the listed A and B patches, computed by
``annf_pca_cov_dev``: the sums are exact integers, so the matrix equals the
CPU one bit for bit.  ``fit_pca_gpu`` finishes with the same LAPACK eigh
and the same sign rule (largest |loading| of each component positive,
sklearn's svd_flip with u_based_decision=False), giving the components and
mean sklearn's PCA(svd_solver="covariance_eigh") fits on those patches.
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


def fit_pca_gpu(image_a, idx_a, psize, d, image_b=None, idx_b=None, device=0, handle=None):
    """(components [d, D], mean [D]) of the listed patches, covariance on the GPU."""
    mean, cov = pca_covariance(image_a, idx_a, image_b, idx_b, psize, device, handle)
    return components_from_covariance(cov, d), mean
