"""Writes the golden fixtures in tests/golden/ from the reference's own openmp
C core (annfmp/openmp/kdtree/src/*.c), compiled by ``make -C oracle ref``
into oracle/_ref/libannf_ref.so and driven through ctypes.

Run in a container that has the reference sources:

    make -C oracle ref && python tests/golden/make_golden.py

The fixtures hold only data: inputs and the reference's outputs.  The
reference itself never travels with this repository.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O          # noqa: E402
from annfmp_amd import synth            # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def _basis(pa, pb, d):
    """A fixed FP64 PCA basis (covariance eigenvectors, descending)."""
    X = np.concatenate([pa, pb]).astype(np.float64)
    mu = X.mean(0)
    w, V = np.linalg.eigh(np.cov(X.T))
    return V[:, ::-1][:, :d].T.copy(), mu


def tree_cases(rng):
    cases = {}
    n = 3001
    cases["gauss"] = (rng.standard_normal((n, 6)).astype(np.float32), 16)
    # heavy ties: few distinct values per coordinate
    cases["ties"] = (rng.integers(0, 4, (2500, 4)).astype(np.float32), 8)
    # duplicated rows and a constant column
    X = rng.standard_normal((1537, 5)).astype(np.float32)
    X[:, 2] = 0.5
    X[700:1400] = X[0:700]
    cases["dups"] = (X, 12)
    # one dimension, signed zeros and repeats
    X = rng.integers(-3, 4, (777, 1)).astype(np.float32)
    X[X == 0] = np.where(rng.random((X == 0).sum()) < 0.5, -0.0, 0.0).astype(np.float32)
    cases["1d"] = (X, 5)
    return cases


def write_trees(rng):
    for name, (X, leaf) in tree_cases(rng).items():
        depth = O.tree_depth(X.shape[0], leaf)
        K = 4
        t = O.RefTree(X, depth, K)
        nq = 64
        if name == "ties":
            Q = rng.integers(0, 4, (nq, X.shape[1])).astype(np.float32)
        elif name == "1d":
            Q = rng.integers(-3, 4, (nq, 1)).astype(np.float32)
        else:
            Q = rng.standard_normal((nq, X.shape[1])).astype(np.float32)
        nat = np.array([t.natural_leaf(q) for q in Q], np.int32)
        ex_d = np.zeros((nq, K), np.float32)
        ex_p = np.zeros((nq, K), np.int32)
        bf_lists = rng.integers(0, 1 << depth, (nq, 1 + K)).astype(np.int32)
        bf_lists[:, 2] = bf_lists[:, 0]          # a leaf listed twice
        bf_d = np.zeros((nq, K), np.float32)
        bf_p = np.zeros((nq, K), np.int32)
        for i, q in enumerate(Q):
            ex_d[i], ex_p[i] = t.exact_knn(q, K)
            bf_d[i], bf_p[i] = t.brute_force_group(q, bf_lists[i], K)
        nb_d, nb_i = t.neighbors(Q, K)
        np.savez_compressed(
            os.path.join(OUT, f"tree_{name}.npz"), X=X, leaf_size=leaf, depth=depth,
            pts=t.pts, orig=t.orig, axis=t.axis, split=t.split, leaves=t.leaves, pli=t.pli,
            K=K, Q=Q, natural=nat, exact_dist=ex_d, exact_pos=ex_p,
            group_leaves=bf_lists, group_dist=bf_d, group_pos=bf_p,
            neighbors_dist=nb_d, neighbors_idx=nb_i)
        print("tree", name, X.shape, depth)


ROW_CASES = [
    # name, H, W, p, d, leaf, K, propagate, select_best_nn, spec kwargs
    ("k1", 40, 56, 5, 6, 8, 1, 1, 1, {}),
    ("k3", 40, 56, 5, 6, 8, 3, 1, 1, {}),
    ("k5_noprop", 48, 40, 4, 5, 16, 5, 0, 1, {}),
    ("k4_nosel", 40, 48, 5, 6, 8, 4, 1, 0, {}),
    ("k2_shift", 36, 36, 3, 4, 6, 2, 1, 1, {"mode": "shift"}),
    # 128x128 at the 256x256 workload's patch/PCA/leaf shape (p=7, d=10, leaf 32)
    # (K=1 and K=5 share one file: final holds K=1, final_k5 the K=5 field)
    ("cfg0_shape", 128, 128, 7, 10, 32, 1, 1, 1, {"mode": "shift", "n_rects": 10, "sigma": 2.0}),
]


def write_rows():
    for name, H, W, p, d, leaf, K, prop, sel, kw in ROW_CASES:
        spec = synth.small_spec(H, W, **kw)
        a, b = synth.make_pair(spec, seed=11)
        pa = O.extract_patches(a, p)
        pb = O.extract_patches(b, p)
        comp, mu = _basis(pa, pb, d)
        ra = O.project_f64(pa, comp, mu)
        rb = O.project_f64(pb, comp, mu)
        nr, nc = H - p + 1, W - p + 1
        depth = O.tree_depth(rb.shape[0], leaf)
        t = O.RefTree(rb, depth, K)
        final = t.process_rows(pa, pb, ra, nr, nc, K, prop, sel)
        extra = {}
        if name == "cfg0_shape":
            t5 = O.RefTree(rb, depth, 5)
            extra["final_k5"] = t5.process_rows(pa, pb, ra, nr, nc, 5, prop, sel)
        np.savez_compressed(
            os.path.join(OUT, f"rows_{name}.npz"), image_a=a, image_b=b, psize=p, red_a=ra, red_b=rb,
            components=comp, mean=mu, leaf_size=leaf, depth=depth, n_rows=nr, n_cols=nc, K=K,
            propagate=prop, select_best_nn=sel, final=final, **extra)
        print("rows", name, final.shape)


def write_norm(rng):
    D = 363
    A = rng.integers(0, 256, (200, D)).astype(np.uint8)
    B = rng.integers(0, 256, (200, D)).astype(np.uint8)
    A[0] = 0; B[0] = 255          # largest distance: FP32 sum rounds past 2**24
    A[1] = 255; B[1] = 0
    A[2] = B[2]                    # zero distance
    out = np.array([O.ref_norm_array(A[i], B[i]) for i in range(A.shape[0])], np.float32)
    np.savez_compressed(os.path.join(OUT, "norm_array.npz"), A=A, B=B, out=out)
    print("norm_array", out.shape)


def main():
    if not O.ref_available():
        raise SystemExit("build the reference core first: make -C oracle ref")
    rng = np.random.default_rng(20201210)
    write_trees(rng)
    write_rows()
    write_norm(rng)


if __name__ == "__main__":
    main()
