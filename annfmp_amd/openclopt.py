"""The openclopt backend's semantics on the GPU.

Mirrors what ``ANNFieldPropKDTreeOpenCLOPT.fit`` (annfmp/openclopt/base.py:47-130)
computes after its PCA fit: reduced patches of both images, then
``fit_extern`` (annfmp/openclopt/src/base.c:257-520), which runs
futhark/driverKNN.fut -- the padded power-of-two tree of buildKDtree.fut, left
on ``<=`` descent, K fixed to 8 (driverKNN.fut:16), the exact 8-NN of row 0 by
the stackless traversal of kd-traverse.fut, the row chain over de-duplicated
propagated leaves, and the best of the eight candidates by full-patch
distance.  The kernels are ``annf_field_openclopt_dev``; the projection is
``annf_project_dev`` (FP32 by default, as the reference's reducePatchDim).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import device as dev

K = 8


def project_on_device(img, psize, components, mean, projection="fp32", handle=None, device=0):
    """(n, d) float32 reduced patches of one image (annf_project_dev)."""
    torch = dev.torch()
    cuda = dev.cuda_device(device)
    h = handle or _lib.Handle(device)
    comp = np.ascontiguousarray(components, dtype=np.float64 if projection == "fp64" else np.float32)
    d = comp.shape[0]
    H, W = img.shape[:2]
    n = (H - psize + 1) * (W - psize + 1)
    ti = torch.from_numpy(np.ascontiguousarray(img, dtype=np.uint8)).to(cuda)
    tc = torch.from_numpy(comp).to(cuda)
    tm = torch.from_numpy(np.ascontiguousarray(mean, dtype=np.float64)).to(cuda)
    out = torch.empty((n, d), dtype=torch.float32, device=cuda)
    proj = _lib.PROJ_PCA64 if projection == "fp64" else _lib.PROJ_PCA32
    _lib.check("annf_project_dev", _lib.lib().annf_project_dev(
        h.h, dev.dptr(ti), H, W, int(psize), dev.dptr(tc), dev.dptr(tm), d, proj, dev.dptr(out)))
    return out.cpu().numpy()


def field_openclopt_on_device(red_a, n_rows, n_cols, red_b, leaf_size, image_a=None, image_b=None, psize=None,
                              handle=None, device=0, return_tree=False):
    """The openclopt pipeline on reduced patches.  Returns a dict: ``knn``
    (n_a x 8 original B indices), ``knd`` (their squared reduced distances),
    ``nn`` (when both images are given: the best of the 8 by exact
    full-patch distance) and, with ``return_tree``, the tree arrays
    (``shape`` = (height, q, ppl, m'), ``dims``, ``vals``, ``clanc``,
    ``indir``, ``orig2leaf``)."""
    torch = dev.torch()
    cuda = dev.cuda_device(device)
    h = handle or _lib.Handle(device)
    L = _lib.lib()
    red_a = np.ascontiguousarray(red_a, dtype=np.float32)
    red_b = np.ascontiguousarray(red_b, dtype=np.float32)
    n_a, d = red_a.shape
    n_b = red_b.shape[0]
    ta = torch.from_numpy(red_a).to(cuda)
    tb = torch.from_numpy(red_b).to(cuda)
    knn = torch.empty((n_a, K), dtype=torch.int32, device=cuda)
    knd = torch.empty((n_a, K), dtype=torch.float32, device=cuda)
    nn = None
    ia = ib = None
    ha = wa = hb = wb = 0
    if image_a is not None:
        ia = torch.from_numpy(np.ascontiguousarray(image_a, dtype=np.uint8)).to(cuda)
        ib = torch.from_numpy(np.ascontiguousarray(image_b, dtype=np.uint8)).to(cuda)
        ha, wa = image_a.shape[:2]
        hb, wb = image_b.shape[:2]
        nn = torch.empty(n_a, dtype=torch.int32, device=cuda)
    _lib.check("annf_field_openclopt_dev", L.annf_field_openclopt_dev(
        h.h, dev.dptr(ta), int(n_rows), int(n_cols), dev.dptr(tb), n_b, d, int(leaf_size), dev.dptr(ia), ha, wa,
        dev.dptr(ib), hb, wb, int(psize or 0), dev.dptr(nn), dev.dptr(knn), dev.dptr(knd)))
    out = {"knn": knn.cpu().numpy(), "knd": knd.cpu().numpy()}
    if nn is not None:
        out["nn"] = nn.cpu().numpy()
    if return_tree:
        out.update(openclopt_tree(h))
    return out


def openclopt_tree(handle):
    """The handle's last openclopt tree (annf_openclopt_tree_export)."""
    L = _lib.lib()
    shape = np.zeros(4, np.int32)
    _lib.check("annf_openclopt_tree_export", L.annf_openclopt_tree_export(
        handle.h, _lib.ptr(shape, _lib._ip), None, None, None, None, None))
    height, q, ppl, mp = (int(x) for x in shape)
    dims = np.zeros(max(q, 1), np.int32)
    vals = np.zeros(max(q, 1), np.float32)
    clanc = np.zeros(max(q, 1), np.int32)
    indir = np.zeros(mp, np.int32)
    o2l = np.zeros(mp, np.int32)
    _lib.check("annf_openclopt_tree_export", L.annf_openclopt_tree_export(
        handle.h, None, _lib.ptr(dims, _lib._ip), _lib.ptr(vals, _lib._fp), _lib.ptr(clanc, _lib._ip),
        _lib.ptr(indir, _lib._ip), _lib.ptr(o2l, _lib._ip)))
    return {"shape": (height, q, ppl, mp), "dims": dims[:q], "vals": vals[:q], "clanc": clanc[:q], "indir": indir,
            "orig2leaf": o2l}


def field_openclopt(image_a, image_b, psize, components, mean, leaf_size, projection="fp32", handle=None, device=0):
    """Image-level openclopt field (base.c:257-520 after the PCA fit):
    project both images, then ``field_openclopt_on_device``."""
    h = handle or _lib.Handle(device)
    red_a = project_on_device(image_a, psize, components, mean, projection, h, device)
    red_b = project_on_device(image_b, psize, components, mean, projection, h, device)
    nr, nc = image_a.shape[0] - psize + 1, image_a.shape[1] - psize + 1
    res = field_openclopt_on_device(red_a, nr, nc, red_b, leaf_size, image_a, image_b, psize, handle=h,
                                    device=device)
    res["red_a"], res["red_b"] = red_a, red_b
    return res
