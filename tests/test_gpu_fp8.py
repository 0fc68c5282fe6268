"""GPU tests of the FP8 full-patch leaf pass's pieces (annf_patch_fp8_dev and
the ANNF_PREC_FP8 form of annf_field_mm_dev).

The per-patch E4M3 codes, inverse scales and norms equal the oracle's
restatement (oracle.patch_fp8) bit for bit.  The FP8 natural-leaf pass is
checked against a float64 emulation on the same codes: per query the
`shortlist` smallest approximate distances d~ = n_a + n_b - 2 (q_a . q_b)
u_a u_b over the probed leaves, re-ranked exactly.  The GPU accumulates the
code dot products in FP32 on MFMA and orders by d~ rounded down to its key
bits, so a query may differ where two approximate distances at the
shortlist's edge are within that rounding; the test bounds how often.
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


def _fp8_emulated(pa, pb, ca, na, ia, cb, nb, ib, qleaf, leaves, orig, nprobe, S, K):
    """Per query: the S smallest float64 approximate distances over the
    probed leaves, then the K smallest exact (d2, index); also the gap
    between the S-th and (S+1)-th approximate distances."""
    va = O.e4m3_value(ca)
    vb = O.e4m3_value(cb)
    lv = leaves.reshape(-1, 2)
    out_i = np.full((len(qleaf), K), -1, np.int64)
    gap = np.full(len(qleaf), np.inf)
    for q in range(len(qleaf)):
        idx = np.concatenate([orig[lv[qleaf[q] ^ r, 0]:lv[qleaf[q] ^ r, 1]] for r in range(nprobe)])
        d = na[q] + nb[idx] - 2.0 * (vb[idx] @ va[q]) * float(ia[q]) * ib[idx]
        order = np.lexsort((idx, d))
        short = idx[order[:S]]
        if len(idx) > S:
            ds = np.sort(d)
            gap[q] = (ds[S] - ds[S - 1]) / max(abs(ds[S]), 1.0)
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
