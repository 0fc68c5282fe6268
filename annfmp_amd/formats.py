"""On-disk formats of a fitted k-d tree and of a nearest-neighbour field.

The reference keeps neither: its tree lives in a KD_TREE_RECORD
(annfmp/openmp/kdtree/src/include/global.h:38-50) for the life of the process
and ``fit`` returns the field as an array.  Both formats here are numpy
``.npz`` archives of plain arrays, read with ``allow_pickle=False`` (nothing in
a file is executed), tagged with a ``format`` name and a ``version``.

tree, ``annfmp_amd.kdtree`` version 1 -- the record's arrays:

    n, d, depth         int64 scalars
    axis, split         int32 / float32, 2^depth - 1 internal nodes (heap order)
    leaves              int32, 2 * 2^depth: [from, to) of every leaf
    xi_orig             int32, n: original index of each tree position
    xi_pts              float32, n x d: the points in tree order
    pli                 int32, n: leaf of each original index
    params              str: the KDTree constructor parameters as JSON

field, ``annfmp_amd.field`` version 1 -- the output of a fit:

    nn                  int32, n_rows * n_cols: ``nn[y * n_cols + x]`` is the
                        original index of the B patch matched to A patch (y, x)
    n_rows, n_cols      int64 scalars (A's patch grid), psize int64
    b_rows, b_cols      int64 scalars (B's patch grid)
    knn                 int32, n x K (optional): the K candidates before best-of-K
    dist                float32, n (optional): L2 of each chosen match
    meta                str: JSON (method, parameters, timings)

``load_*`` checks the tag, the version, the shapes and the index ranges, and
raises ``FormatError`` on any mismatch.
"""
from __future__ import annotations

import json

import numpy as np

TREE_FORMAT = "annfmp_amd.kdtree"
FIELD_FORMAT = "annfmp_amd.field"
VERSION = 1


class FormatError(ValueError):
    pass


def _open(path, fmt):
    try:
        z = np.load(path, allow_pickle=False)
    except ValueError as e:   # object arrays are refused, never unpickled
        raise FormatError("{}: not a plain-array archive ({})".format(path, e))
    if not hasattr(z, "files"):
        raise FormatError("{}: not an .npz archive".format(path))
    files = set(z.files)
    if "format" not in files or "version" not in files:
        raise FormatError("{}: no format tag".format(path))
    tag = str(z["format"])
    if tag != fmt:
        raise FormatError("{}: format {!r}, expected {!r}".format(path, tag, fmt))
    ver = int(z["version"])
    if ver != VERSION:
        raise FormatError("{}: {} version {} (this build reads version {})".format(path, fmt, ver, VERSION))
    return z


def _need(z, path, names):
    missing = [k for k in names if k not in z.files]
    if missing:
        raise FormatError("{}: missing {}".format(path, ", ".join(missing)))


def _arr(z, name, dtype, shape, path):
    a = z[name]
    if a.dtype != np.dtype(dtype) or a.shape != tuple(shape):
        raise FormatError("{}: {} is {} {}, expected {} {}".format(path, name, a.dtype, a.shape,
                                                                 np.dtype(dtype), tuple(shape)))
    return np.ascontiguousarray(a)


# ---------------------------------------------------------------------------
def save_tree(path, arrays, params=None):
    """Writes the arrays of ``KDTree.tree_arrays()`` (plus the constructor
    parameters) to ``path``."""
    n, d = arrays["xi_pts"].shape
    np.savez(path, format=np.array(TREE_FORMAT), version=np.array(VERSION, np.int64),
             n=np.array(n, np.int64), d=np.array(d, np.int64), depth=np.array(int(arrays["depth"]), np.int64),
             axis=np.asarray(arrays["axis"], np.int32), split=np.asarray(arrays["split"], np.float32),
             leaves=np.asarray(arrays["leaves"], np.int32), xi_orig=np.asarray(arrays["xi_orig"], np.int32),
             xi_pts=np.asarray(arrays["xi_pts"], np.float32), pli=np.asarray(arrays["pli"], np.int32),
             params=np.array(json.dumps(params or {}, sort_keys=True)))


def load_tree(path):
    """Reads a tree file: dict of the arrays, ``depth``, ``n``, ``d`` and
    ``params`` (dict).  The leaf ranges, permutation and point-to-leaf map
    are checked again by ``annf_tree_import`` before a tree is used."""
    z = _open(path, TREE_FORMAT)
    _need(z, path, ("n", "d", "depth", "axis", "split", "leaves", "xi_orig", "xi_pts", "pli", "params"))
    n, d, depth = int(z["n"]), int(z["d"]), int(z["depth"])
    if n < 1 or d < 1 or not 0 <= depth <= 30 or (n >> depth) < 1:
        raise FormatError("{}: n={} d={} depth={} is not a valid tree".format(path, n, d, depth))
    nn = (1 << depth) - 1
    out = {
        "n": n, "d": d, "depth": depth,
        "axis": _arr(z, "axis", np.int32, (nn,), path),
        "split": _arr(z, "split", np.float32, (nn,), path),
        "leaves": _arr(z, "leaves", np.int32, (2 << depth,), path),
        "xi_orig": _arr(z, "xi_orig", np.int32, (n,), path),
        "xi_pts": _arr(z, "xi_pts", np.float32, (n, d), path),
        "pli": _arr(z, "pli", np.int32, (n,), path),
        "params": json.loads(str(z["params"])),
    }
    if nn and (out["axis"].min() < 0 or out["axis"].max() >= d):
        raise FormatError("{}: split axis outside [0, {})".format(path, d))
    if not np.array_equal(np.sort(out["xi_orig"]), np.arange(n, dtype=np.int32)):
        raise FormatError("{}: xi_orig is not a permutation of [0, {})".format(path, n))
    return out


# ---------------------------------------------------------------------------
def save_field(path, nn, n_rows, n_cols, psize, b_shape, knn=None, dist=None, meta=None):
    """Writes a field: ``nn`` over A's n_rows x n_cols patch grid, indices
    into B's patch grid of ``b_shape`` = (b_rows, b_cols)."""
    nn = np.asarray(nn, np.int32).ravel()
    if nn.shape[0] != n_rows * n_cols:
        raise FormatError("nn has {} entries, the grid {} x {}".format(nn.shape[0], n_rows, n_cols))
    extra = {}
    if knn is not None:
        extra["knn"] = np.asarray(knn, np.int32).reshape(nn.shape[0], -1)
    if dist is not None:
        extra["dist"] = np.asarray(dist, np.float32).ravel()
    np.savez(path, format=np.array(FIELD_FORMAT), version=np.array(VERSION, np.int64), nn=nn,
             n_rows=np.array(n_rows, np.int64), n_cols=np.array(n_cols, np.int64),
             psize=np.array(psize, np.int64), b_rows=np.array(b_shape[0], np.int64),
             b_cols=np.array(b_shape[1], np.int64), meta=np.array(json.dumps(meta or {}, sort_keys=True)),
             **extra)


def load_field(path):
    """Reads a field file: dict with nn, n_rows, n_cols, psize, b_rows,
    b_cols, knn / dist (None when absent) and meta (dict)."""
    z = _open(path, FIELD_FORMAT)
    _need(z, path, ("nn", "n_rows", "n_cols", "psize", "b_rows", "b_cols", "meta"))
    n_rows, n_cols = int(z["n_rows"]), int(z["n_cols"])
    b_rows, b_cols = int(z["b_rows"]), int(z["b_cols"])
    n = n_rows * n_cols
    out = {"n_rows": n_rows, "n_cols": n_cols, "psize": int(z["psize"]), "b_rows": b_rows, "b_cols": b_cols,
           "nn": _arr(z, "nn", np.int32, (n,), path), "knn": None, "dist": None,
           "meta": json.loads(str(z["meta"]))}
    nb = b_rows * b_cols
    if n and (out["nn"].min() < 0 or out["nn"].max() >= nb):
        raise FormatError("{}: nn indexes outside B's {} patches".format(path, nb))
    if "knn" in z.files:
        knn = z["knn"]
        if knn.dtype != np.int32 or knn.ndim != 2 or knn.shape[0] != n:
            raise FormatError("{}: knn is {} {}, expected int32 ({}, K)".format(path, knn.dtype, knn.shape, n))
        if knn.size and (knn.min() < -1 or knn.max() >= nb):   # -1: an empty slot
            raise FormatError("{}: knn indexes outside B's {} patches".format(path, nb))
        out["knn"] = np.ascontiguousarray(knn)
    if "dist" in z.files:
        out["dist"] = _arr(z, "dist", np.float32, (n,), path)
    return out
