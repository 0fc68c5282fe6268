"""Throughput of the nearest-neighbour-field hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--nprobe 1]
                    [--encoder annfmp_amd/data/encoder_p11_d16.npz]

Workload: the 1920x1080 synthetic warp pair (seed 0), p = 11 (363-byte
patches), d = 16, leaf 64 (depth 14), K = 5.  A step is one whole field,
from device-resident uint8 images to the int32 field, for one of two paths:

* learned + BF16 (the measured `value`): fused extraction + MLP encoder
  (363-128-128-16, BF16 on MFMA) for A and B, tree build over B's
  embeddings, natural leaf of every A patch, full-patch k-NN over that leaf
  on MFMA (BF16 operands, exact integer distances).  PCA-16 replaces the
  encoder when no trained encoder is given or when the encoder's recall@5
  falls below PCA's on this pair (the fallback);
* plain FP32 (reported beside it): the reference's method -- PCA-16
  projection in FP32, tree, row 0 exact search, natural-leaf scan,
  propagation sweep, best-of-K -- bit-exact to the openmp backend.

W untimed steps, then K timed steps bracketed by device synchronisation,
for each path.  Quality (recall@5 and mean distance ratio against the
exact field on 8192 sampled A patches) is measured outside the timed
region.  The PCA basis is fitted once on 50,000 sampled A patches (FP64
covariance and eigensolve on the GPU).
Prints one JSON line.
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
BF16_PEAK_TFS = 2500.0         # MI355X dense BF16 MFMA, 2.5 PFLOP/s
# stage boundaries on the main stream: B is projected first; A is projected on
# a second stream during the tree build, whose stage ends once A is done too
PLAIN_STAGES = ["project_b", "launch_project_a", "tree_build_with_project_a", "natural_leaf_scan", "row0_exact",
                "row_sweep", "best_of_k"]
MM_STAGES = ["project_b", "launch_project_a", "tree_build_with_project_a", "natural_leaf_bucket", "leaf_knn_mfma",
             "propagated_pass"]
DEFAULT_ENCODER = os.path.join(ROOT, "annfmp_amd", "data", "encoder_p11_d16.npz")
DEFAULT_ENCODER_4K = os.path.join(ROOT, "annfmp_amd", "data", "encoder_p9_d16.npz")
PATHS_DEFAULT = "plain,pca,learned"
FP8_PEAK_TFS = 5000.0          # MI355X dense FP8 MFMA, 5.0 PFLOP/s
FP32_PEAK_TFS = 157.3          # MI355X FP32, 157.3 TFLOP/s


def fit_basis(a, p, d, n_samples, seed):
    """PCA basis from n_samples uniformly sampled A patches: FP64 covariance
    and eigensolve on the GPU (annf_pca_fit_dev)."""
    from annfmp_amd.encoder import fit_pca
    return fit_pca(a, p, d, n_samples, seed)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class _quiet_stdout:
    """Sends C-level stdout to /dev/null (the reference core prints its
    progress), so that this script's one JSON line stays alone."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def cpu_baseline(a, b, p, red_a, red_b, depth, K, gpu_nn):
    """The plain method on the host cores on the same float32 reduced arrays
    (tree build + row 0, propagation, conversion and best-of-K): the
    reference's own C core (oracle/_ref, compiled from its sources) when it
    is built, else the C restatement; on every host core of this process and
    on one thread."""
    from oracle import oracle as O
    cores = len(os.sched_getaffinity(0))
    pa = O.extract_patches(a, p)
    pb = O.extract_patches(b, p)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    n = red_a.shape[0]
    kind = "reference" if O.ref_available() else "port"

    def run(threads):
        t0 = time.time()
        if kind == "reference":
            with _quiet_stdout():
                ref = O.RefTree(red_b, depth, K=K, num_threads=threads)
                t1 = time.time()
                final = ref.process_rows(pa, pb, red_a, nr, nc, K, True, True)
        else:
            ref = O.OracleTree(red_b, depth)
            t1 = time.time()
            final, _, _ = ref.process_rows(red_a, nr, nc, K, True, True, pa, pb, nthreads=threads)
        t2 = time.time()
        return final, t1 - t0, t2 - t1

    f_all, fit_all, rows_all = run(cores)
    f_one, fit_one, rows_one = run(1)
    parity = bool(np.array_equal(f_all, gpu_nn) and np.array_equal(f_one, gpu_nn))
    return {
        "value": round(n / (fit_all + rows_all), 1), "unit": "query patches/s", "cores": cores, "kind": kind,
        "value_1thread": round(n / (fit_one + rows_one), 1), "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
        "sample": ("plain path on the full 1920x1080 pair, %d queries, same float32 reduced arrays as the GPU run: "
                   "%s; %d threads: kdtree_fit %.2f s (serial, as the reference) + rows/conversion/best-of-K "
                   "%.2f s; 1 thread: %.2f s + %.2f s"
                   % (n, "the reference's own openmp C core (fit_extern + process_rows_extern)"
                      if kind == "reference" else "the C restatement of the openmp core",
                      cores, fit_all, rows_all, fit_one, rows_one)),
    }, parity


def path_roofline(p, d, n, n_leaf_pts, ms_per_step, learned):
    """Path-wide roofline of the measured path: t_roof = sum over stages of
    max(algorithmic bytes / HBM peak, algorithmic flops / MFMA peak) per
    query, against the measured time per query.  Stages: projection of A
    and B (encoder MLP 363-128-128-16 on BF16 MFMA, or PCA on FP32), the
    tree build over B, natural leaf + bucketing, the BF16 full-patch leaf
    k-NN over one leaf."""
    D = p * p * 3
    bw = HBM_PEAK_GBS * 1e9
    img = 2 * 3.0   # two images, 3 bytes per pixel, about one pixel per patch
    if learned:
        s1_flops, s1_peak = 2 * 2 * (D * 128 + 128 * 128 + 128 * d), BF16_PEAK_TFS * 1e12
    else:
        s1_flops, s1_peak = 2 * 2 * D * d, FP32_PEAK_TFS * 1e12
    stages = {
        "project_a_b": (img + 2 * 4 * d, s1_flops, s1_peak),
        "tree_build": (4 * d + (4 * d + 4), 0.0, 1.0),
        "natural_leaf_bucket": (4 * d + 4, 0.0, 1.0),
        "leaf_knn_bf16": (img + 4 + 4 + 4 * 5, 2.0 * D * n_leaf_pts, BF16_PEAK_TFS * 1e12),
    }
    t = {k: max(bts / bw, fl / pk) for k, (bts, fl, pk) in stages.items()}
    t_roof = sum(t.values())
    t_meas = ms_per_step * 1e-3 / n
    return {"t_roof_ps_per_query": round(t_roof * 1e12, 2), "t_measured_ps_per_query": round(t_meas * 1e12, 2),
            "frac": round(t_roof / t_meas, 4), "ceiling_query_patches_per_s": round(1.0 / t_roof, 1),
            "stages_ps": {k: round(v * 1e12, 2) for k, v in t.items()},
            "peaks": {"hbm_GBs": HBM_PEAK_GBS, "bf16_TFs": BF16_PEAK_TFS, "fp32_TFs": FP32_PEAK_TFS}}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), or
    None when the summary has no line for it."""
    import re
    path = os.path.join(ROOT, "profiles", "leaf_knn_bf16_1080p_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch " + re.escape(kernel) + r": (\d+)", open(path).read())
    return int(m.group(1)) if m else None


def timed(fn, steps, warmup, torch, lib, h, nst):
    """Mean ms per step over `steps` timed calls and the mean stage times."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = np.zeros(8, np.float32)
    acc = np.zeros(nst)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        lib.lib().annf_stage_times(h.h, lib.ptr(st, lib._fp), nst)
        acc += st[:nst]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1000.0 / steps, acc / steps


def run_4k(args):
    """The combined form at 4K: 3840x2160 warp pair, p = 9 (243-dim), d = 16,
    leaf 128 (depth 15), K = 8.  Times the plain FP32 path (PCA-16,
    propagation, best-of-K) and the learned encoder with FP8 full-patch leaf
    distances (top-32 exact re-rank) and the propagated pass, plus the same
    with BF16; recall@8 against the exact field.  One JSON line."""
    import torch
    from annfmp_amd import synth, _lib, metrics, device as dev
    from annfmp_amd.kdtree import tree_depth_for
    from annfmp_amd.encoder import PatchEncoder

    p, d, leaf, K, S = 9, 16, 128, 8, 32
    prop = K if args.propagate is None else args.propagate
    a, b = synth.make_pair("4k-warp", seed=args.seed)
    H, W, _ = a.shape
    comp, mu = fit_basis(a, p, d, 50000, args.seed)
    n_a = (H - p + 1) * (W - p + 1)
    n_b = n_a
    depth = tree_depth_for(n_b, leaf)
    cuda = dev.cuda_device(0)
    ia, ib = torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda)
    tc = torch.from_numpy(comp.astype(np.float32)).to(cuda)
    tm = torch.from_numpy(mu.astype(np.float64)).to(cuda)
    nn = torch.empty(n_a, dtype=torch.int32, device=cuda)
    knn = torch.empty((n_a, K), dtype=torch.int32, device=cuda)
    h = _lib.Handle(0)
    L = _lib.lib()
    _lib.check("annf_set_timing", L.annf_set_timing(h.h, 1))
    enc_path = args.encoder if args.encoder != DEFAULT_ENCODER else DEFAULT_ENCODER_4K
    encoder = None
    if args.encoder != "none" and os.path.exists(enc_path):
        encoder = PatchEncoder.load(enc_path)
        encoder.install(h)
    proj = _lib.PROJ_ENCODER if encoder is not None else _lib.PROJ_PCA32

    def plain(knn_out=None):
        _lib.check("annf_field_dev", L.annf_field_dev(
            h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d, _lib.PROJ_PCA32,
            depth, K, 1, 1, dev.dptr(nn), None, dev.dptr(knn_out), None, None))

    def mm(prec, shortlist, propagate, knn_out=None):
        _lib.check("annf_field_mm_dev", L.annf_field_mm_dev(
            h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d, proj, depth, K, 1,
            prec, shortlist, propagate, dev.dptr(nn), None, dev.dptr(knn_out), None, None))

    if not args.no_quality:
        q = metrics.sample_queries(n_a, 8192)
        exact = metrics.exact_knn(a, b, p, q, K, handle=h)

    def quality():
        torch.cuda.synchronize()
        if args.no_quality:
            return {"recall@8": float("nan"), "distance_ratio": float("nan")}
        return metrics.field_quality(a, b, p, nn.cpu().numpy(), knn.cpu().numpy(), k=K, handle=h, exact=exact)

    # --paths other than the 1080p default selects runs by name (plain, learned_fp8_prop, ...)
    select = None if args.paths == PATHS_DEFAULT else set(args.paths.split(","))
    paths = {}
    if select is None or "plain" in select:
        ms, stage = timed(plain, args.steps, args.warmup, torch, _lib, h, 7)
        plain(knn)
        paths["plain_fp32"] = {"ms_per_step": round(ms, 3), "query_patches_per_s": round(n_a / ms * 1e3, 1),
                               "stage_ms": {s_: round(float(v), 3) for s_, v in zip(PLAIN_STAGES, stage)},
                               **quality()}
    flops_pass1 = 2.0 * (p * p * 3) * n_a * (n_b / (1 << depth))
    runs = (("learned_fp8_prop", _lib.PREC_FP8, S, prop), ("learned_bf16_prop", _lib.PREC_BF16, K, prop),
            ("learned_bf16_prop2", _lib.PREC_BF16, K, 2), ("learned_fp8", _lib.PREC_FP8, S, 0),
            ("learned_bf16", _lib.PREC_BF16, K, 0))
    for name, prec, sl, pr in runs:
        if select is not None and name not in select:
            continue
        ms, stage = timed(lambda: mm(prec, sl, pr), args.steps, args.warmup, torch, _lib, h, 6)
        mm(prec, sl, pr, knn)
        peak = FP8_PEAK_TFS if prec == _lib.PREC_FP8 else BF16_PEAK_TFS
        leaf_ms = float(stage[4])
        paths[name] = {"ms_per_step": round(ms, 3), "query_patches_per_s": round(n_a / ms * 1e3, 1),
                       "precision": "fp8" if prec == _lib.PREC_FP8 else "bf16", "shortlist": sl,
                       "propagated_entries": pr,
                       "stage_ms": {s_: round(float(v), 3) for s_, v in zip(MM_STAGES, stage)},
                       "leaf_pass_mfma_tflops": round(flops_pass1 / (leaf_ms * 1e-3) / 1e12, 2),
                       "leaf_pass_mfma_util": round(flops_pass1 / (leaf_ms * 1e-3) / 1e12 / peak, 5),
                       **quality()}
    if "learned_fp8_prop" not in paths or "plain_fp32" not in paths:
        print(json.dumps({"paths": paths}), flush=True)
        return
    main_p = paths["learned_fp8_prop"]
    out = {
        "metric": "query patches/sec, 4K pair: learned+FP8 (top-32 re-rank, propagation) vs plain FP32, recall@8",
        "value": main_p["query_patches_per_s"], "unit": "query patches/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": main_p["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp8", "data": "synthetic",
        "config": {"workload": "3840x2160 warp pair (seed %d), p=9 (243-dim), %s, leaf 128 (depth %d), K=8, "
                               "FP8-E4M3 full-patch leaf distances, top-32 exact re-rank, %d propagated entries"
                               % (args.seed, "learned encoder 243-128-128-16" if encoder is not None else "PCA-16",
                                  depth, prop),
                   "queries_per_step": n_a, "encoder": os.path.relpath(enc_path, ROOT) if encoder else None},
        "speedup_vs_plain_fp32": round(main_p["query_patches_per_s"] / paths["plain_fp32"]["query_patches_per_s"], 3),
        "paths": paths,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nprobe", type=int, default=1, help="leaves probed per query on the BF16 path")
    ap.add_argument("--encoder", default=DEFAULT_ENCODER, help="trained encoder (.npz); 'none' forces PCA")
    ap.add_argument("--paths", default=PATHS_DEFAULT,
                    help="paths to run (profiling runs select one); the value needs pca and learned; "
                         "learned_prop / learned_chain add the learned BF16 form with the parallel propagated "
                         "pass / the row chain")
    ap.add_argument("--no-quality", action="store_true", help="skip the exact-field quality evaluation")
    ap.add_argument("--workload", default="1080p", choices=["1080p", "4k"],
                    help="1080p: the metric (default); 4k: the combined form on the 3840x2160 pair")
    ap.add_argument("--propagate", type=int, default=None,
                    help="4k: list entries that propagate on the MFMA path (default K = 8)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run 5 device-to-device copies of MB MiB (scales the HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.empty(args.copy_calibration << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration << 20, "copies": 5}))
        return
    run_paths = set(args.paths.split(","))
    if args.workload == "4k":
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        run_4k(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the field would shard by image columns; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    from annfmp_amd import synth, _lib, metrics, device as dev
    from annfmp_amd.kdtree import tree_depth_for
    from annfmp_amd.encoder import PatchEncoder

    p, d, leaf, K = 11, 16, 64, 5
    a, b = synth.make_pair("1080p-warp", seed=args.seed)
    H, W, _ = a.shape
    comp, mu = fit_basis(a, p, d, 50000, args.seed)
    n_a = (H - p + 1) * (W - p + 1)
    n_b = n_a
    depth = tree_depth_for(n_b, leaf)

    cuda = dev.cuda_device(0)
    ia = torch.from_numpy(a).to(cuda)
    ib = torch.from_numpy(b).to(cuda)
    tc = torch.from_numpy(comp.astype(np.float32)).to(cuda)
    tm = torch.from_numpy(mu.astype(np.float64)).to(cuda)
    nn = torch.empty(n_a, dtype=torch.int32, device=cuda)
    knn = torch.empty((n_a, K), dtype=torch.int32, device=cuda)
    ra = torch.empty((n_a, d), dtype=torch.float32, device=cuda)
    rb = torch.empty((n_b, d), dtype=torch.float32, device=cuda)
    h = _lib.Handle(0)
    L = _lib.lib()
    _lib.check("annf_set_timing", L.annf_set_timing(h.h, 1))

    encoder = None
    if args.encoder != "none" and os.path.exists(args.encoder):
        encoder = PatchEncoder.load(args.encoder)
        encoder.install(h)

    def plain(knn_out=None, red=False):
        _lib.check("annf_field_dev", L.annf_field_dev(
            h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d, _lib.PROJ_PCA32,
            depth, K, 1, 1, dev.dptr(nn), None, dev.dptr(knn_out), dev.dptr(ra) if red else None,
            dev.dptr(rb) if red else None))

    def mm(proj, knn_out=None):
        _lib.check("annf_field_mm_dev", L.annf_field_mm_dev(
            h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d, proj, depth, K,
            args.nprobe, _lib.PREC_BF16, K, 0, dev.dptr(nn), None, dev.dptr(knn_out), None, None))

    if not args.no_quality:
        q = metrics.sample_queries(n_a, 8192)
        exact = metrics.exact_knn(a, b, p, q, 5, handle=h)

    def quality():
        torch.cuda.synchronize()
        if args.no_quality:
            return {"recall@5": float("nan"), "distance_ratio": float("nan")}
        return metrics.field_quality(a, b, p, nn.cpu().numpy(), knn.cpu().numpy(), k=5, handle=h, exact=exact)

    paths = {}
    plain_nn = None
    if "plain" in run_paths:
        # plain FP32 reference method
        ms, stage = timed(plain, args.steps, args.warmup, torch, _lib, h, 7)
        plain(knn, red=True)
        plain_nn = nn.cpu().numpy()
        paths["plain_fp32"] = {"ms_per_step": round(ms, 4), "query_patches_per_s": round(n_a / ms * 1e3, 1),
                               "stage_ms": {s: round(float(v), 4) for s, v in zip(PLAIN_STAGES, stage)},
                               **quality()}
    if "pca" in run_paths:
        # PCA + BF16 full-patch leaf k-NN
        ms, stage = timed(lambda: mm(_lib.PROJ_PCA32), args.steps, args.warmup, torch, _lib, h, 6)
        mm(_lib.PROJ_PCA32, knn)
        paths["pca_bf16"] = {"ms_per_step": round(ms, 4), "query_patches_per_s": round(n_a / ms * 1e3, 1),
                             "stage_ms": {s: round(float(v), 4) for s, v in zip(MM_STAGES, stage)}, **quality()}
    chosen = "pca_bf16" if "pca_bf16" in paths else None
    if encoder is not None and "learned" in run_paths:
        ms, stage = timed(lambda: mm(_lib.PROJ_ENCODER), args.steps, args.warmup, torch, _lib, h, 6)
        mm(_lib.PROJ_ENCODER, knn)
        paths["learned_bf16"] = {"ms_per_step": round(ms, 4), "query_patches_per_s": round(n_a / ms * 1e3, 1),
                                 "stage_ms": {s: round(float(v), 4) for s, v in zip(MM_STAGES, stage)},
                                 **quality()}
        if chosen is None or args.no_quality or \
                paths["learned_bf16"]["recall@5"] >= paths["pca_bf16"]["recall@5"]:
            chosen = "learned_bf16"
    if encoder is not None and ("learned_prop" in run_paths or "learned_chain" in run_paths):
        # the propagated pass beside the metric's path: one parallel pass from
        # the first-pass lists of the row above, or the reference's row chain
        # (rows in order from the row above's final lists); not the value
        for name, prop in (("learned_prop", K), ("learned_chain", -K)):
            if name not in run_paths:
                continue
            def mm_prop(knn_out=None, prop=prop):
                _lib.check("annf_field_mm_dev", L.annf_field_mm_dev(
                    h.h, dev.dptr(ia), H, W, dev.dptr(ib), H, W, p, dev.dptr(tc), dev.dptr(tm), d,
                    _lib.PROJ_ENCODER, depth, K, args.nprobe, _lib.PREC_BF16, K, prop, dev.dptr(nn), None,
                    dev.dptr(knn_out), None, None))
            ms, stage = timed(mm_prop, max(1, args.steps // 5), 1, torch, _lib, h, 6)
            mm_prop(knn)
            paths[name] = {"ms_per_step": round(ms, 4), "query_patches_per_s": round(n_a / ms * 1e3, 1),
                           "stage_ms": {s: round(float(v), 4) for s, v in zip(MM_STAGES, stage)}, **quality()}
    if chosen is None:
        print(json.dumps({"paths": paths}), flush=True)
        return
    main_path = paths[chosen]
    ms_per_step = main_path["ms_per_step"]
    value = n_a / ms_per_step * 1e3

    # roofline of the dominant kernel of the measured path: the MFMA leaf
    # k-NN (stage marks around its one launch); algorithmic work = 2 D flops
    # per (query, leaf point) pair of the probed leaves
    leaf_ms = main_path["stage_ms"]["leaf_knn_mfma"]
    flops = 2.0 * (p * p * 3) * n_a * args.nprobe * (n_b / (1 << depth))
    achieved = flops / (leaf_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "k_leaf_bf16<28, 5, true> (BF16)", "achieved": round(achieved, 3),
                "peak": BF16_PEAK_TFS, "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFS, 6),
                "traffic": measured_traffic("k_leaf_bf16<28, 5, true>"), "algorithmic_flops": flops,
                "avg_ms": round(float(leaf_ms), 4)}
    path_roof = path_roofline(p, d, n_a, args.nprobe * n_b / (1 << depth), ms_per_step, chosen == "learned_bf16")

    cpu = None
    parity = None
    if not args.no_cpu and plain_nn is not None:
        cpu, parity = cpu_baseline(a, b, p, ra.cpu().numpy(), rb.cpu().numpy(), depth, K, plain_nn)

    plain_res = paths.get("plain_fp32", {"recall@5": float("nan"), "query_patches_per_s": float("nan")})
    plain_rec = plain_res["recall@5"]
    out = {
        "metric": "query patches/sec, 1080p pair: learned+BF16 vs plain FP32 at equal recall@5",
        "value": round(value, 1), "unit": "query patches/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": ("1920x1080 warp pair (seed %d), p=11 (363-dim), %s + BF16 full-patch leaf k-NN on "
                                "MFMA, leaf 64 (depth 14), %d leaf probed, K=5; plain FP32 path (PCA-16, "
                                "propagation + best-of-K) timed beside it")
                               % (args.seed, "learned encoder 363-128-128-16" if chosen == "learned_bf16"
                                  else "PCA-16", args.nprobe),
                   "queries_per_step": n_a, "seed": args.seed, "path": chosen,
                   "encoder": os.path.relpath(args.encoder, ROOT) if encoder is not None else None},
        "speedup_vs_plain_fp32": round(value / plain_res["query_patches_per_s"], 3),
        "recall@5_vs_plain_fp32": [round(main_path["recall@5"], 4), round(plain_rec, 4)],
        "recall_at_least_plain": bool(main_path["recall@5"] >= plain_rec),
        "paths": paths, "roofline": roofline, "path_roofline": path_roof, "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
