"""Times the C restatement (liboracle.so) against the reference's own C core
(oracle/_ref/libannf_ref.so, compiled from the reference's sources by
oracle/Makefile) on the same float32 reduced arrays: tree build
(fit_extern) and the rows (process_rows_extern: row 0, propagation,
conversion, best-of-K), at 1 thread and at every host core.  Both must give
the same field.  The ratio justifies timing the restatement where the
reference core is not built.

    python oracle/time_vs_reference.py [--workloads 256,1080p] [--threads 1,8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def workload(name):
    from annfmp_amd import synth
    from annfmp_amd.encoder import fit_pca
    if name == "256":
        a, b = synth.make_pair("256x256-shift", seed=0)
        p, d, leaf, K, ns = 7, 10, 32, 1, 10000
    else:
        a, b = synth.make_pair("1080p-warp", seed=0)
        p, d, leaf, K, ns = 11, 16, 64, 5, 50000
    comp, mu = fit_pca(a, p, d, ns, seed=0, where="host")
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    ra, rb = O.project_f64(pa, comp, mu), O.project_f64(pb, comp, mu)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    return pa, pb, ra, rb, nr, nc, O.tree_depth(pb.shape[0], leaf), K


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="256,1080p")
    ap.add_argument("--threads", default="1,%d" % len(os.sched_getaffinity(0)))
    args = ap.parse_args()
    if not O.ref_available():
        sys.exit("the reference core is not built (make -C oracle ref)")
    out = {"cpu": cpu_model(), "host_cores": len(os.sched_getaffinity(0)), "runs": []}
    for w in args.workloads.split(","):
        pa, pb, ra, rb, nr, nc, depth, K = workload(w)
        for th in [int(x) for x in args.threads.split(",")]:
            t0 = time.perf_counter()
            mine = O.OracleTree(rb, depth)
            t1 = time.perf_counter()
            f_mine, _, _ = mine.process_rows(ra, nr, nc, K, True, True, pa, pb, nthreads=th)
            t2 = time.perf_counter()
            ref = O.RefTree(rb, depth, K=K, num_threads=th)
            t3 = time.perf_counter()
            f_ref = ref.process_rows(pa, pb, ra, nr, nc, K, True, True)
            t4 = time.perf_counter()
            run = {"workload": w, "threads": th, "queries": int(ra.shape[0]),
                   "restatement_s": {"fit": round(t1 - t0, 3), "rows": round(t2 - t1, 3)},
                   "reference_s": {"fit": round(t3 - t2, 3), "rows": round(t4 - t3, 3)},
                   "ratio_restatement_over_reference": round((t2 - t0) / (t4 - t2), 3),
                   "same_field": bool(np.array_equal(f_mine, f_ref))}
            print(json.dumps(run), flush=True)
            out["runs"].append(run)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
