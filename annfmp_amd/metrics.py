"""Quality of a nearest-neighbour field against the exact field.

Definitions (distances are L2 over the full p x p x 3 uint8 patch):

* exact field -- the k nearest B patches of an A patch by exact squared L2,
  ties broken by the lower B index (``annf_exact_field_dev``, integer
  arithmetic, identical to an FP64 brute force);
* recall@k -- for each sampled A patch, the fraction of the method's k
  candidates (its K list, duplicates counted once, padded with misses) whose
  distance is at most the exact k-th nearest distance; averaged over the
  sample.  Counting by distance rather than by index makes ties at the k-th
  distance count as hits;
* distance ratio -- mean over the sample of d(method's chosen match) /
  d(exact nearest), over patches whose exact nearest distance is non-zero;
* score -- the mean L2 distance of every A patch to its chosen match, the
  quality number the reference's example prints (examples/simple.py:31-36).

Queries are a seeded uniform sample of A patches; the exact field is
computed on the GPU for the sample only.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from . import device as dev


def sample_queries(n_a, n_sample, seed=12345):
    rng = np.random.default_rng(seed)
    if n_sample >= n_a:
        return np.arange(n_a, dtype=np.int32)
    return np.sort(rng.choice(n_a, size=n_sample, replace=False)).astype(np.int32)


def exact_knn(image_a, image_b, p, query_idx, k, device=0, handle=None, dev_images=None):
    """Exact k nearest B patches of the listed A patches: (idx, d2) int32."""
    torch = dev.torch()
    cuda = dev.cuda_device(device)
    h = handle or _lib.Handle(device)
    if dev_images is None:
        ia = torch.from_numpy(np.ascontiguousarray(image_a, dtype=np.uint8)).to(cuda)
        ib = torch.from_numpy(np.ascontiguousarray(image_b, dtype=np.uint8)).to(cuda)
    else:
        ia, ib = dev_images
    ha, wa = image_a.shape[:2]
    hb, wb = image_b.shape[:2]
    q = torch.from_numpy(np.ascontiguousarray(query_idx, dtype=np.int32)).to(cuda)
    nq = int(q.numel())
    idx = torch.empty((nq, k), dtype=torch.int32, device=cuda)
    d2 = torch.empty((nq, k), dtype=torch.int32, device=cuda)
    torch.cuda.synchronize(cuda)
    _lib.check("annf_exact_field_dev", _lib.lib().annf_exact_field_dev(
        h.h, dev.dptr(ia), ha, wa, dev.dptr(ib), hb, wb, int(p), dev.dptr(q), nq, int(k), dev.dptr(idx),
        dev.dptr(d2)))
    return idx.cpu().numpy(), d2.cpu().numpy()


def patch_d2(image_a, image_b, p, a_idx, b_idx):
    """Exact squared L2 between A patches a_idx and B patches b_idx (host)."""
    a_idx = np.asarray(a_idx, np.int64).ravel()
    b_idx = np.asarray(b_idx, np.int64).ravel()
    nca = image_a.shape[1] - p + 1
    ncb = image_b.shape[1] - p + 1
    ya, xa = a_idx // nca, a_idx % nca
    yb, xb = b_idx // ncb, b_idx % ncb
    out = np.zeros(a_idx.shape[0], np.int64)
    A = image_a.astype(np.int32)
    B = image_b.astype(np.int32)
    for r in range(p):
        for c in range(p):
            diff = A[ya + r, xa + c, :] - B[yb + r, xb + c, :]
            out += (diff * diff).sum(1)
    return out


def field_quality(image_a, image_b, p, nn, knn, k=5, n_sample=8192, seed=12345, device=0, handle=None,
                  exact=None):
    """recall@k of the K lists, distance ratio of the chosen matches, score.

    nn: (N,) chosen B index per A patch; knn: (N, K) candidate lists (may be
    None, then nn alone is the list).  Returns a dict."""
    n_a = nn.shape[0]
    q = sample_queries(n_a, n_sample, seed)
    if exact is None:
        e_idx, e_d2 = exact_knn(image_a, image_b, p, q, k, device=device, handle=handle)
    else:
        e_idx, e_d2 = exact
    lists = knn[q] if knn is not None else nn[q][:, None]
    kk = min(k, lists.shape[1])
    S, W = lists.shape
    # the first k distinct candidates of every list
    distinct = np.ones((S, W), bool)
    for j in range(1, W):
        distinct[:, j] = np.all(lists[:, :j] != lists[:, j:j + 1], axis=1)
    distinct &= np.cumsum(distinct, axis=1) <= k
    d2 = patch_d2(image_a, image_b, p, np.repeat(q, W), lists.ravel()).reshape(S, W)
    hits = ((d2 <= e_d2[:, k - 1:k].astype(np.int64)) & distinct).sum(1) / k
    chosen = patch_d2(image_a, image_b, p, q, nn[q])
    best = e_d2[:, 0].astype(np.int64)
    nz = best > 0
    ratio = float(np.mean(np.sqrt(chosen[nz]) / np.sqrt(best[nz]))) if nz.any() else 1.0
    return {"recall@%d" % k: float(hits.mean()), "distance_ratio": ratio, "list_width": int(kk),
            "sample": int(q.shape[0]), "exact_nn_zero": int((~nz).sum())}


def score(dist):
    """Mean L2 of the chosen matches (examples/simple.py:31-36)."""
    return float(np.mean(dist))
