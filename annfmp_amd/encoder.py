"""Learned patch encoder: an MLP D -> 128 -> 128 -> d replacing PCA as the
k-d tree's projection, trained on the GPU with the library's own HIP kernels.

Training data are triplets (query patch, its exact nearest B patch, a
uniformly random B patch) drawn from seeded synthetic pairs; the exact
nearest patch comes from the library's brute-force field
(``annf_exact_field_dev``).  The loss is the triplet margin loss (margin
0.1), optimised with Adam.  Forward, backward and the optimiser step are the
kernels behind ``annf_mlp_triplet_grad_dev`` and ``annf_mlp_adam_dev``;
PyTorch only holds the device buffers.  Inference is the fused
extraction + MLP kernel behind ``annf_set_encoder`` / ANNF_PROJ_ENCODER.

Weights are saved as ``.npz`` (plain arrays, loaded with allow_pickle off).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib, metrics, synth
from . import device as dev

H1 = 128
H2 = 128


def param_count(D, h1, h2, d):
    return h1 * D + h1 + h2 * h1 + h2 + d * h2 + d


def init_params(D, h1, h2, d, seed=0):
    """Linear-layer default initialisation: weights and biases uniform in
    +-1/sqrt(fan_in), drawn layer by layer (weights, then bias)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((D, h1), (h1, h2), (h2, d)):
        bound = 1.0 / np.sqrt(fan_in)
        parts.append(rng.uniform(-bound, bound, fan_out * fan_in))
        parts.append(rng.uniform(-bound, bound, fan_out))
    return np.concatenate(parts).astype(np.float32)


class PatchEncoder:
    """Parameters of the MLP for patch size p (D = p*p*3 inputs)."""

    def __init__(self, psize, d=16, h1=H1, h2=H2, params=None, seed=0):
        self.psize, self.d, self.h1, self.h2 = int(psize), int(d), int(h1), int(h2)
        self.D = self.psize * self.psize * 3
        n = param_count(self.D, self.h1, self.h2, self.d)
        if params is None:
            params = init_params(self.D, self.h1, self.h2, self.d, seed)
        params = np.ascontiguousarray(params, dtype=np.float32)
        if params.size != n:
            raise ValueError(f"{params.size} parameters given, {n} expected")
        self.params = params
        self.history = []

    def save(self, path, **meta):
        np.savez(path, params=self.params, psize=self.psize, d=self.d, h1=self.h1, h2=self.h2,
                 history=np.asarray(self.history, np.float64), **meta)

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        enc = cls(int(z["psize"]), int(z["d"]), int(z["h1"]), int(z["h2"]), params=z["params"])
        enc.history = z["history"].tolist() if "history" in z else []
        return enc

    def install(self, handle, device=0):
        """Copies the parameters into the handle's fused encoder."""
        torch = dev.torch()
        t = torch.from_numpy(self.params).to(dev.cuda_device(device))
        _lib.check("annf_set_encoder", _lib.lib().annf_set_encoder(handle.h, dev.dptr(t), self.psize, self.h1,
                                                                   self.h2, self.d))
        torch.cuda.synchronize()

    def encode_image(self, image, handle=None, device=0):
        """Fused extraction + encoding of every patch of a host image."""
        torch = dev.torch()
        h = handle or _lib.Handle(device)
        self.install(h, device)
        cuda = dev.cuda_device(device)
        img = torch.from_numpy(np.ascontiguousarray(image, dtype=np.uint8)).to(cuda)
        H, W = image.shape[:2]
        n = (H - self.psize + 1) * (W - self.psize + 1)
        out = torch.empty((n, self.d), dtype=torch.float32, device=cuda)
        _lib.check("annf_encode_image_dev", _lib.lib().annf_encode_image_dev(h.h, dev.dptr(img), H, W, self.psize,
                                                                             dev.dptr(out)))
        return out.cpu().numpy()


def patch_rows(img, p, idx):
    """uint8 rows (len(idx), p*p*3) of the listed patches (a3 order)."""
    H, W, _ = img.shape
    nc = W - p + 1
    idx = np.asarray(idx, np.int64)
    ys, xs = idx // nc, idx % nc
    win = np.lib.stride_tricks.sliding_window_view(img, (p, p, 3))[:, :, 0]
    return np.ascontiguousarray(win[ys, xs].reshape(len(idx), p * p * 3))


def make_triplets(p, n_pairs, per_pair, spec="512x512-warp", seed=1000, handle=None, device=0, log=None):
    """(n_pairs * per_pair, 3, D) uint8 triplets: query A patch, its exact
    nearest B patch (ties to the lower index), a uniform random B patch.
    Pair s is synth.make_pair(spec, seed + s)."""
    h = handle or _lib.Handle(device)
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_pairs):
        a, b = synth.make_pair(spec, seed=seed + s)
        n_a = (a.shape[0] - p + 1) * (a.shape[1] - p + 1)
        n_b = (b.shape[0] - p + 1) * (b.shape[1] - p + 1)
        q = np.sort(rng.choice(n_a, size=per_pair, replace=False)).astype(np.int32)
        idx, _ = metrics.exact_knn(a, b, p, q, 1, device=device, handle=h)
        neg = rng.integers(0, n_b, per_pair)
        out.append(np.stack([patch_rows(a, p, q), patch_rows(b, p, idx[:, 0]), patch_rows(b, p, neg)], 1))
        if log:
            log(f"triplets: pair {s + 1}/{n_pairs}")
    return np.concatenate(out)


def train(encoder, triplets, epochs=2, batch=1024, lr=1e-3, margin=0.1, beta1=0.9, beta2=0.999, eps=1e-8, seed=0,
          handle=None, device=0, log=None, log_every=200):
    """Adam on the triplet margin loss; updates encoder.params in place and
    appends (step, loss) pairs to encoder.history."""
    torch = dev.torch()
    h = handle or _lib.Handle(device)
    L = _lib.lib()
    cuda = dev.cuda_device(device)
    D, h1, h2, d = encoder.D, encoder.h1, encoder.h2, encoder.d
    if triplets.shape[1:] != (3, D):
        raise ValueError(f"triplets of shape {triplets.shape}, expected (T, 3, {D})")
    n = encoder.params.size
    params = torch.from_numpy(encoder.params.copy()).to(cuda)
    grads = torch.zeros(n, dtype=torch.float32, device=cuda)
    m = torch.zeros(n, dtype=torch.float32, device=cuda)
    v = torch.zeros(n, dtype=torch.float32, device=cuda)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    rng = np.random.default_rng(seed)
    T = triplets.shape[0]
    nb = T // batch
    if nb < 1:
        raise ValueError("fewer triplets than one batch")
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        perm = rng.permutation(T)[:nb * batch]
        # batches laid out as [queries; positives; negatives] rows
        host = np.ascontiguousarray(triplets[perm].reshape(nb, batch, 3, D).transpose(0, 2, 1, 3))
        xb = torch.from_numpy(host).to(cuda)
        torch.cuda.synchronize()
        for i in range(nb):
            _lib.check("annf_mlp_triplet_grad_dev", L.annf_mlp_triplet_grad_dev(
                h.h, dev.dptr(xb[i]), batch, D, dev.dptr(params), h1, h2, d, float(margin), dev.dptr(grads),
                dev.dptr(loss)))
            step += 1
            _lib.check("annf_mlp_adam_dev", L.annf_mlp_adam_dev(
                h.h, dev.dptr(params), dev.dptr(grads), dev.dptr(m), dev.dptr(v), n, float(lr), float(beta1),
                float(beta2), float(eps), step))
            if step % log_every == 0 or i == nb - 1:
                lv = float(loss.item())
                encoder.history.append((step, lv))
                if log:
                    log(f"epoch {ep} step {step} loss {lv:.5f} ({time.time() - t0:.1f} s)")
        del xb
    encoder.params = params.cpu().numpy()
    return encoder


def fit_pca(a, p, d, n_samples=50000, seed=0, where="gpu", handle=None, device=0):
    """PCA basis (d x D) and mean from n_samples uniformly sampled A patches
    (positions from default_rng(seed + 1000)): FP64 covariance and
    eigensolve, sklearn's sign rule (largest |loading| of each component
    positive).  where="gpu": the whole fit on the GPU (annf_pca_fit_dev);
    "host": numpy on the host (for CPU-only scripts)."""
    rng = np.random.default_rng(seed + 1000)
    H, W, _ = a.shape
    ys = rng.integers(0, H - p + 1, n_samples)
    xs = rng.integers(0, W - p + 1, n_samples)
    if where == "gpu":
        from .pca import fit_pca_gpu
        idx = (ys * (W - p + 1) + xs).astype(np.int32)
        return fit_pca_gpu(a, idx, p, d, handle=handle, device=device)
    if where != "host":
        raise ValueError("where must be 'gpu' or 'host'")
    X = np.stack([a[y:y + p, x:x + p, :].reshape(-1) for y, x in zip(ys, xs)]).astype(np.float64)
    mu = X.mean(0)
    C = (X - mu).T @ (X - mu) / (n_samples - 1)
    _, V = np.linalg.eigh(C)
    comp = V[:, ::-1][:, :d].T.copy()
    comp *= np.sign(comp[np.arange(d), np.abs(comp).argmax(1)])[:, None]
    return comp, mu


def compare_with_pca(encoder, seeds, spec="1080p-warp", leaf=64, K=5, nprobe=1, n_sample=8192, handle=None,
                     device=0, log=None):
    """recall@5 and mean distance ratio of the BF16 full-patch leaf k-NN with
    the encoder's tree against PCA-d's tree on the same pairs, at equal leaf
    visits (nprobe leaves of the same size per query).  Returns a list of
    per-pair dicts."""
    from .base import field_mm_on_device
    from .kdtree import tree_depth_for
    h = handle or _lib.Handle(device)
    p = encoder.psize
    out = []
    for s in seeds:
        a, b = synth.make_pair(spec, seed=s)
        comp, mu = fit_pca(a, p, encoder.d, seed=s, handle=h)
        n_a = (a.shape[0] - p + 1) * (a.shape[1] - p + 1)
        n_b = (b.shape[0] - p + 1) * (b.shape[1] - p + 1)
        depth = tree_depth_for(n_b, leaf)
        q = metrics.sample_queries(n_a, n_sample)
        exact = metrics.exact_knn(a, b, p, q, 5, device=device, handle=h)
        row = {"seed": s}
        for name, proj in (("pca", "fp32"), ("encoder", encoder)):
            r = field_mm_on_device(a, b, p, comp, mu, depth, K, nprobe=nprobe, precision="bf16", projection=proj,
                                   handle=h, device=device)
            qual = metrics.field_quality(a, b, p, r["nn"], r["knn"], k=5, n_sample=n_sample, handle=h, exact=exact)
            row[name] = {"recall@5": qual["recall@5"], "distance_ratio": qual["distance_ratio"]}
        out.append(row)
        if log:
            log(f"pair seed {s}: {row}")
    return out


def train_default(p=11, d=16, n_pairs=64, per_pair=31250, epochs=2, batch=1024, lr=1e-3, seed=1000, handle=None,
                  device=0, log=None):
    """The encoder of the learned workload: 64 seeded 512x512 pairs, 2M
    triplets, margin 0.1.  Returns (encoder, seconds for triplets, seconds
    for training)."""
    h = handle or _lib.Handle(device)
    t0 = time.time()
    trip = make_triplets(p, n_pairs, per_pair, seed=seed, handle=h, device=device, log=log)
    t1 = time.time()
    enc = PatchEncoder(p, d, seed=seed)
    train(enc, trip, epochs=epochs, batch=batch, lr=lr, seed=seed, handle=h, device=device, log=log)
    return enc, t1 - t0, time.time() - t1


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the learned patch encoder on the GPU")
    ap.add_argument("--psize", type=int, default=11)
    ap.add_argument("--d", type=int, default=16)
    ap.add_argument("--pairs", type=int, default=64, help="seeded 512x512 training pairs")
    ap.add_argument("--per-pair", type=int, default=31250, help="triplets per pair (64 x 31250 = 2M)")
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000, help="first training pair seed (and init seed)")
    ap.add_argument("--eval-seeds", default="0,1,2,3", help="held-out 1080p pairs for the PCA comparison")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    h = _lib.Handle(0)
    enc, t_trip, t_train = train_default(args.psize, args.d, args.pairs, args.per_pair, args.epochs, args.batch,
                                         args.lr, args.seed, handle=h, log=log)
    seeds = [int(x) for x in args.eval_seeds.split(",") if x != ""]
    cmp = compare_with_pca(enc, seeds, handle=h, log=log) if seeds else []
    enc.save(args.out, triplet_seconds=t_trip, train_seconds=t_train, pairs=args.pairs, per_pair=args.per_pair,
             epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed,
             eval_json=json.dumps(cmp))
    log(json.dumps({"out": args.out, "triplet_s": round(t_trip, 2), "train_s": round(t_train, 2), "eval": cmp}))


if __name__ == "__main__":
    main()
