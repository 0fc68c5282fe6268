"""Evaluation CLI for nearest-neighbour fields.

Reports, for one image pair, the quality of a field against the exact field
(``metrics``): recall@k of the K candidate lists, the mean distance ratio of
the chosen matches, and the mean L2 score of every A patch to its match, the
number examples/simple.py:31-36 prints.  The field is either computed here
with the HIP backend or read from a field file (``formats``)::

    python -m annfmp_amd.evaluate A.png B.png --method learned-bf16 --psize 11 --out field.npz
    python -m annfmp_amd.evaluate A.png B.png --field field.npz
    python -m annfmp_amd.evaluate --synthetic 1080p-warp --seed 0 --method plain --psize 11 --dim 16

Images are ``.npy`` (H x W x 3 uint8) or any file PIL reads (converted to
RGB).  Methods:

    plain         the reference's method: PCA, tree, reduced-space leaf
                  search, row propagation, best-of-K (ANNFieldPropKDTreeOpenMP)
    pca-bf16      PCA tree, full-patch leaf distances in BF16 on MFMA
    pca-fp8       PCA tree, FP8-E4M3 leaf distances + exact re-rank
    learned-bf16  the learned encoder's tree, BF16 leaf distances
    learned-fp8   the learned encoder's tree, FP8 leaf distances

The exact field is computed on the GPU for a seeded sample of A patches
(``--sample``); the score covers every A patch.  One JSON line goes to
stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METHODS = ("plain", "pca-bf16", "pca-fp8", "learned-bf16", "learned-fp8")
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_image(path):
    """uint8 (H, W, 3) array from a .npy file or an image file."""
    if path.endswith(".npy"):
        img = np.load(path, allow_pickle=False)
    else:
        from PIL import Image
        with Image.open(path) as im:
            img = np.asarray(im.convert("RGB"))
    img = np.ascontiguousarray(img)
    if img.dtype != np.uint8 or img.ndim != 3 or img.shape[2] != 3:
        raise ValueError("{}: expected an H x W x 3 uint8 image, got {} {}".format(path, img.dtype, img.shape))
    return img


def default_encoder(psize, dim):
    path = os.path.join(DATA, "encoder_p{}_d{}.npz".format(psize, dim))
    return path if os.path.exists(path) else None


def full_patch_l2(image_a, image_b, p, nn, chunk=1 << 18):
    """L2 over the full p x p x 3 patch of every A patch to nn[i] (float32)."""
    from .metrics import patch_d2
    n = nn.shape[0]
    out = np.empty(n, np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        out[s:e] = np.sqrt(patch_d2(image_a, image_b, p, np.arange(s, e), nn[s:e]).astype(np.float64))
    return out


def compute_field(image_a, image_b, method, args):
    """Runs the HIP backend; returns (nn, knn, dist, timers)."""
    from .base import ANNFieldPropKDTreeHIP
    kw = dict(n_neighbors=args.neighbors, psize=args.psize, dim_reduced=args.dim, n_subset=args.n_subset,
              leaf_size=args.leaf, propagate=not args.no_propagate, select_best_nn=True, float_type="float",
              verbose=0, seed=args.seed, device=args.device, projection=args.projection)
    if method != "plain":
        kw["leaf_distance"] = method.split("-")[1]
        kw["shortlist"] = args.shortlist
        kw["nprobe"] = args.nprobe
    if method.startswith("learned"):
        enc = args.encoder or default_encoder(args.psize, args.dim)
        if enc is None:
            raise SystemExit("no encoder for psize {} dim {}: pass --encoder".format(args.psize, args.dim))
        kw["encoder"] = enc
    model = ANNFieldPropKDTreeHIP(**kw)
    t0 = time.time()
    nn = model.fit(image_a, image_b)
    wall = time.time() - t0
    timers = {k: float(v) for k, v in model._timers.items()}
    timers["fit_wall_s"] = wall
    return nn, model.knn_indices, model.nn_distances, timers


def evaluate(image_a, image_b, psize, nn, knn=None, dist=None, k=5, sample=8192, seed=12345, device=0):
    """Quality dict of a field: recall@k, distance_ratio, score."""
    from .metrics import field_quality, score
    q = field_quality(image_a, image_b, psize, nn, knn, k=k, n_sample=sample, seed=seed, device=device)
    if dist is None:
        dist = full_patch_l2(image_a, image_b, psize, nn)
    q["score"] = score(dist)
    return q


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m annfmp_amd.evaluate", description=__doc__.split("\n\n")[0])
    ap.add_argument("images", nargs="*", help="image A and image B")
    ap.add_argument("--synthetic", help="a seeded synthetic pair instead of images (256x256-shift, 512x512-warp, "
                                        "1080p-warp, 4k-warp)")
    ap.add_argument("--seed", type=int, default=0, help="pair seed (synthetic) and PCA sampling seed")
    ap.add_argument("--method", choices=METHODS, default="plain")
    ap.add_argument("--field", help="evaluate this field file instead of computing one")
    ap.add_argument("--out", help="write the computed field to this file")
    ap.add_argument("--psize", type=int, default=8)
    ap.add_argument("--dim", type=int, default=16, help="dim_reduced (PCA components or encoder outputs)")
    ap.add_argument("--leaf", type=int, default=64, help="leaf_size")
    ap.add_argument("--neighbors", type=int, default=5, help="n_neighbors (K)")
    ap.add_argument("--n-subset", type=int, default=1000, help="patches sampled for the PCA fit")
    ap.add_argument("--projection", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--encoder", help="encoder .npz for the learned methods")
    ap.add_argument("--shortlist", type=int, default=20, help="FP8 re-rank length")
    ap.add_argument("--nprobe", type=int, default=1)
    ap.add_argument("--no-propagate", action="store_true")
    ap.add_argument("--k", type=int, default=None, help="recall@k (default: n_neighbors)")
    ap.add_argument("--sample", type=int, default=8192, help="A patches checked against the exact field")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)

    if args.synthetic:
        if args.images:
            ap.error("give either two images or --synthetic")
        from . import synth
        image_a, image_b = synth.make_pair(args.synthetic, seed=args.seed)
        pair = "synthetic {} seed {}".format(args.synthetic, args.seed)
    else:
        if len(args.images) != 2:
            ap.error("two images (A and B) are needed")
        image_a, image_b = load_image(args.images[0]), load_image(args.images[1])
        pair = "{} -> {}".format(*args.images)

    from . import device as dev
    dev.cuda_device(args.device)   # the GPU is initialised before the library loads
    out = {"pair": pair, "shape_a": list(image_a.shape), "shape_b": list(image_b.shape)}
    if args.field:
        from . import formats
        f = formats.load_field(args.field)
        p = f["psize"]
        if (f["n_rows"], f["n_cols"]) != (image_a.shape[0] - p + 1, image_a.shape[1] - p + 1) or \
                (f["b_rows"], f["b_cols"]) != (image_b.shape[0] - p + 1, image_b.shape[1] - p + 1):
            raise SystemExit("{}: the field's patch grids do not match the images at psize {}".format(args.field, p))
        nn, knn, dist = f["nn"], f["knn"], f["dist"]
        out.update({"field": args.field, "psize": p, "meta": f["meta"]})
        k = args.k or (knn.shape[1] if knn is not None else 1)
    else:
        p = args.psize
        nn, knn, dist, timers = compute_field(image_a, image_b, args.method, args)
        k = args.k or args.neighbors
        out.update({"method": args.method, "psize": p, "dim": args.dim, "leaf": args.leaf,
                    "neighbors": args.neighbors, "timers_s": timers})
        if args.out:
            from . import formats
            meta = {k_: out[k_] for k_ in ("pair", "method", "psize", "dim", "leaf", "neighbors")}
            formats.save_field(args.out, nn, image_a.shape[0] - p + 1, image_a.shape[1] - p + 1, p,
                               (image_b.shape[0] - p + 1, image_b.shape[1] - p + 1), knn=knn, dist=dist, meta=meta)
            out["out"] = args.out
    out.update(evaluate(image_a, image_b, p, nn, knn, dist, k=k, sample=args.sample, device=args.device))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
