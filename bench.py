"""Throughput of the nearest-neighbour-field hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one full field on the 1920x1080 synthetic pair (p = 11 -> 363-dim
patches, PCA-16, leaf 64, K = 5, propagation on, best-of-K on): fused patch
extraction + FP32 projection of both images, tree build over B, row 0 exact
search, natural-leaf scan, propagation sweep, index conversion and best-of-K,
from device-resident uint8 images to the int32 field.  The PCA basis is
fitted once on 50,000 sampled A patches (FP64 covariance, outside the timed
region).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
STAGES = ["project_a", "project_b", "tree_build", "natural_leaf_scan", "row0_exact", "row_sweep", "best_of_k"]
KERNEL_OF_STAGE = {"project_a": "k_project", "project_b": "k_project", "row0_exact": "k_exact",
                   "row_sweep": "k_sweep", "best_of_k": "k_best", "natural_leaf_scan": "k_leafscan"}


def fit_basis(a, p, d, n_samples, seed):
    """PCA basis from n_samples uniformly sampled A patches (FP64 covariance)."""
    rng = np.random.default_rng(seed + 1000)
    H, W, _ = a.shape
    ys = rng.integers(0, H - p + 1, n_samples)
    xs = rng.integers(0, W - p + 1, n_samples)
    X = np.stack([a[y:y + p, x:x + p, :].reshape(-1) for y, x in zip(ys, xs)]).astype(np.float64)
    mu = X.mean(0)
    C = (X - mu).T @ (X - mu) / (n_samples - 1)
    w, V = np.linalg.eigh(C)
    comp = V[:, ::-1][:, :d].T.copy()
    # sklearn's svd_flip sign rule: largest |loading| positive
    sgn = np.sign(comp[np.arange(d), np.abs(comp).argmax(1)])
    comp *= sgn[:, None]
    return comp, mu


def algorithmic_bytes(stage, n_a, n_b, H, W, p, d, dpad, K, ppl):
    """Compulsory bytes of one launch of a stage's kernel (each input read
    once, each output written once)."""
    if stage in ("project_a", "project_b"):
        n = n_a if stage == "project_a" else n_b
        return H * W * 3 + n * dpad * 4 + (p * p * 3) * (dpad + 1) * 4
    if stage == "row_sweep":
        # per query: its vector, the natural-leaf list (dist, pos, leaf),
        # the K propagated leaves' points with their leaf links once each
        # (first touch), and the list written back
        return n_a * (4 * dpad + 12 * K + 4 * K) + n_b * (4 * dpad + 4)
    if stage == "natural_leaf_scan":
        return n_a * (4 * dpad + 4 + 12 * K) + n_b * (4 * dpad)
    if stage == "best_of_k":
        return n_a * (4 * K + 4 + 4) + n_b * 4 + 2 * H * W * 3
    if stage == "row0_exact":
        return 0
    return n_b * (4 * dpad + 4) * 2


def cpu_baseline(a, b, p, comp, mu, red_a, red_b, depth, K, gpu_nn):
    """The oracle (C restatement of the openmp backend, OpenMP) on the same
    float32 reduced arrays: tree build + rows + conversion + best-of-K."""
    from oracle import oracle as O
    threads = min(len(os.sched_getaffinity(0)), 16)
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    t0 = time.time()
    ref = O.OracleTree(red_b, depth)
    t1 = time.time()
    final, _, _ = ref.process_rows(red_a, nr, nc, K, True, True, pa, pb, nthreads=threads)
    t2 = time.time()
    n = red_a.shape[0]
    return {
        "value": n / (t2 - t0), "unit": "query patches/s", "cores": threads, "kind": "port",
        "sample": f"full 1920x1080 pair, {n} queries: kdtree_fit {t1 - t0:.2f} s (serial, as the reference) + "
                  f"rows/conversion/best-of-K {t2 - t1:.2f} s on {threads} threads, same float32 reduced arrays",
    }, bool(np.array_equal(final, gpu_nn))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the field would shard by image columns; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    from annfmp_amd import synth, _lib, device as dev
    from annfmp_amd.kdtree import tree_depth_for

    p, d, leaf, K = 11, 16, 64, 5
    a, b = synth.make_pair("1080p-warp", seed=args.seed)
    H, W, _ = a.shape
    comp, mu = fit_basis(a, p, d, 50000, args.seed)
    n_a = (H - p + 1) * (W - p + 1)
    n_b = n_a
    depth = tree_depth_for(n_b, leaf)
    dpad = (d + 3) // 4 * 4

    cuda = dev.cuda_device(0)
    ia = torch.from_numpy(a).to(cuda)
    ib = torch.from_numpy(b).to(cuda)
    tc = torch.from_numpy(comp.astype(np.float32)).to(cuda)
    tm = torch.from_numpy(mu.astype(np.float64)).to(cuda)
    nn = torch.empty(n_a, dtype=torch.int32, device=cuda)
    ra = torch.empty((n_a, d), dtype=torch.float32, device=cuda)
    rb = torch.empty((n_b, d), dtype=torch.float32, device=cuda)
    h = _lib.Handle(0)
    L = _lib.lib()
    _lib.check("annf_set_timing", L.annf_set_timing(h.h, 1))

    def step(red=False):
        _lib.check("annf_field_dev", L.annf_field_dev(
            h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d, 0, depth, K, 1, 1,
            dev.dptr(nn), None, None, dev.dptr(ra) if red else None, dev.dptr(rb) if red else None))

    torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    stage_sum = np.zeros(7)
    st = np.zeros(7, np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        L.annf_stage_times(h.h, _lib.ptr(st, _lib._fp), 7)
        stage_sum += st
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = n_a * args.steps / (t1 - t0)
    stage_ms = stage_sum / args.steps

    # one more step that also returns the reduced arrays and K lists
    knn = torch.empty((n_a, K), dtype=torch.int32, device=cuda)
    _lib.check("annf_field_dev", L.annf_field_dev(
        h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d, 0, depth, K, 1, 1,
        dev.dptr(nn), None, dev.dptr(knn), dev.dptr(ra), dev.dptr(rb)))
    torch.cuda.synchronize()
    gpu_nn = nn.cpu().numpy()
    from annfmp_amd import metrics
    tq = time.perf_counter()
    quality = metrics.field_quality(a, b, p, gpu_nn, knn.cpu().numpy(), k=5, n_sample=8192, handle=h)
    quality["eval_s"] = round(time.perf_counter() - tq, 2)

    dom = int(np.argmax([stage_ms[i] if STAGES[i] != "tree_build" else 0 for i in range(7)]))
    dom_name = STAGES[dom]
    ppl = n_b >> depth
    abytes = algorithmic_bytes(dom_name, n_a, n_b, H, W, p, d, dpad, K, ppl)
    achieved = abytes / (stage_ms[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": KERNEL_OF_STAGE.get(dom_name, dom_name), "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "algorithmic_bytes": int(abytes), "avg_ms": round(float(stage_ms[dom]), 4)}

    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(a, b, p, comp, mu, ra.cpu().numpy(), rb.cpu().numpy(), depth, K, gpu_nn)

    out = {
        "metric": "query patches/sec, 1080p pair: learned+BF16 vs plain FP32 at equal recall@5",
        "value": round(value, 1), "unit": "query patches/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain FP32 path: 1920x1080 warp pair, p=11 (363-dim), PCA-16 FP32 projection, "
                               "leaf 64 (depth 14), K=5, propagation + best-of-K",
                   "queries_per_step": n_a, "seed": args.seed},
        "stage_ms": {s: round(float(v), 4) for s, v in zip(STAGES, stage_ms)},
        "roofline": roofline, "cpu_baseline": cpu, "parity_vs_oracle": parity, "quality": quality,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
