"""GPU tests of the FP8 full-patch leaf pass's pieces (annf_patch_fp8_dev and
the ANNF_PREC_FP8 form of annf_field_mm_dev).

The per-patch E4M3 codes, inverse scales and norms equal the oracle's
restatement (oracle.patch_fp8) bit for bit.  The FP8 natural-leaf pass is
checked against a float64 emulation on the same codes: per (query, leaf) the
two halves' smallest approximate distances d~ = n_a + n_b - 2 (q_a . q_b)
u_a u_b, the `shortlist` smallest over the probed leaves, re-ranked exactly.
The GPU accumulates the code dot products in FP32 on MFMA and orders by d~
rounded down to its key bits, so a query may differ where two approximate
distances at a cut are within that rounding; the test bounds how often.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p,seed", [(3, 0), (5, 1), (9, 2), (11, 3), (12, 4)])
def test_patch_codes_match_oracle(p, seed, gpu_available):
    from annfmp_amd import synth
    from annfmp_amd.base import patch_fp8_on_device
    a, _ = synth.make_pair(synth.small_spec(40, 56), seed=seed)
    a[5:25, 10:40] = 128          # flat patches: m = 0, scale and inverse 1, codes 0
    a[30:38, 0:20] = 200
    codes, nrm, inv = patch_fp8_on_device(a, p)
    wc, wn, wi = O.patch_fp8(a, p)
    assert codes.shape == wc.shape
    assert np.array_equal(inv.view(np.uint32), wi.view(np.uint32))
    assert np.array_equal(codes, wc)
    assert np.array_equal(nrm.view(np.uint32), wn.view(np.uint32))
    assert (inv == 1.0).any() and (codes.reshape(len(inv), -1)[inv == 1.0] == 0).all()


def _lane_lists(idx, d, S):
    """One (query, leaf) list of the FP8 pass: the leaf's points in ascending
    B index, split by position bit 2 (the two lanes of the MFMA block); the
    LM / 2 smallest of each half (LM = S rounded up to 8, 16 or 32), then the
    S smallest of both.  Returns the kept indices and the gap between the
    last kept and first dropped approximate distance of each half."""
    lm = 8 if S <= 8 else (16 if S <= 16 else 32)
    o = np.argsort(idx, kind="stable")
    idx, d = idx[o], d[o]
    pos = np.arange(len(idx))
    keep_i, keep_d, gap = [], [], np.inf
    for h in (0, 1):
        m = ((pos >> 2) & 1) == h
        hi, hd = idx[m], d[m]
        oo = np.lexsort((hi, hd))
        keep_i.append(hi[oo[:lm // 2]])
        keep_d.append(hd[oo[:lm // 2]])
        if len(oo) > lm // 2:
            ds = hd[oo]
            gap = min(gap, (ds[lm // 2] - ds[lm // 2 - 1]) / max(abs(ds[lm // 2]), 1.0))
    ki, kd = np.concatenate(keep_i), np.concatenate(keep_d)
    oo = np.lexsort((ki, kd))
    if len(oo) > S:
        gap = min(gap, (kd[oo[S]] - kd[oo[S - 1]]) / max(abs(kd[oo[S]]), 1.0))
    return ki[oo[:S]], kd[oo[:S]], gap


def _fp8_emulated(pa, pb, ca, na, ia, cb, nb, ib, qleaf, leaves, orig, nprobe, S, K):
    """Per query: each probed leaf's list (_lane_lists) from float64
    approximate distances, the S smallest over the leaves, then the K
    smallest exact (d2, index); also the smallest gap at any cut."""
    va = O.e4m3_value(ca)
    vb = O.e4m3_value(cb)
    lv = leaves.reshape(-1, 2)
    out_i = np.full((len(qleaf), K), -1, np.int64)
    gap = np.full(len(qleaf), np.inf)
    for q in range(len(qleaf)):
        cand_i, cand_d = [], []
        for r in range(nprobe):
            idx = orig[lv[qleaf[q] ^ r, 0]:lv[qleaf[q] ^ r, 1]]
            d = na[q] + nb[idx] - 2.0 * (vb[idx] @ va[q]) * float(ia[q]) * ib[idx]
            ki, kd, g = _lane_lists(idx, d, S)
            cand_i.append(ki)
            cand_d.append(kd)
            gap[q] = min(gap[q], g)
        ci, cd = np.concatenate(cand_i), np.concatenate(cand_d)
        oo = np.lexsort((ci, cd))
        if len(oo) > S:
            gap[q] = min(gap[q], (cd[oo[S]] - cd[oo[S - 1]]) / max(abs(cd[oo[S]]), 1.0))
        short = ci[oo[:S]]
        d2 = O.exact_pair_d2(pa, pb, np.full(len(short), q), short)
        o2 = np.lexsort((short, d2))[:K]
        out_i[q, :len(o2)] = short[o2]
    return out_i, gap


@pytest.mark.parametrize("p,K,S,nprobe,seed", [(9, 8, 32, 1, 50), (7, 5, 20, 2, 51), (11, 5, 8, 1, 52)])
def test_fp8_natural_pass_matches_emulation(p, K, S, nprobe, seed, gpu_available):
    from annfmp_amd import synth, field_mm_on_device
    from annfmp_amd.base import patch_fp8_on_device
    a, b = synth.make_pair(synth.small_spec(72, 96), seed=seed)
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    X = np.concatenate([pa[::5], pb[::5]]).astype(np.float64)
    mu = X.mean(0)
    _, V = np.linalg.eigh(np.cov(X.T))
    comp = V[:, ::-1][:, :8].T.copy()
    depth = O.tree_depth(pb.shape[0], 32)
    res = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=nprobe, precision="fp8", shortlist=S,
                             return_reduced=True)
    tree = O.OracleTree(res["red_b"], depth)
    qleaf = tree.natural_leaves(res["red_a"])
    ca, na, ia = patch_fp8_on_device(a, p)
    cb, nb, ib = patch_fp8_on_device(b, p)
    want, gap = _fp8_emulated(pa, pb, ca, na.astype(np.float64), ia, cb, nb.astype(np.float64),
                              ib.astype(np.float64), qleaf, tree.leaves, tree.orig, nprobe, S, K)
    same = (res["knn"] == want).all(axis=1)
    print(f"fp8 p={p} S={S}: {same.mean():.5f} of {len(same)} queries equal the emulation")
    # a differing query has its shortlist edge within the arithmetic's reach
    assert (gap[~same] < 1e-3).all()
    assert same.mean() >= 0.99
