"""KDTree on MI355X: the interface of annfmp/openmp/kdtree/base.py.

Same constructor, ``fit``, ``kneighbors``, ``process_rows``, ``get_params``
and tree-depth rule as the reference's SWIG-backed class; the native calls go
to libannf_hip.so (``init_extern`` -> ``annf_init``, ``fit_extern`` ->
``annf_fit``, ``neighbors_extern`` -> ``annf_neighbors``,
``process_rows_extern`` -> ``annf_process_rows``,
``free_resources_extern`` -> ``annf_free_resources``).

Differences, all explicit: only ``float_type="float"`` exists (the reference's
``"double"`` build is refused with an exception); parameter errors raise
``ANNFError`` instead of terminating the process; n_neighbors is limited to
``MAX_K`` and the dimension to ``MAX_D``.
"""
from __future__ import annotations

import math
import multiprocessing
import warnings

import numpy as np

from . import _lib
from ._lib import ANNFError, Handle, check, lib, ptr


class KDTree(object):
    """Nearest neighbor queries based on k-d trees (kdtree/base.py:18-316)."""

    SPLITTING_TYPE_MAPPINGS = {'cyclic': 0, 'longest': 1}
    ALLOWED_FLOAT_TYPES = ['float', 'double']

    def __init__(self, n_neighbors=5, leaf_size=30, max_leaves=None, tree_depth=None, float_type="double",
                 splitting_type="cyclic", n_jobs=1, verbose=0, device=0):
        self.n_neighbors = n_neighbors
        self.leaf_size = leaf_size
        self.max_leaves = max_leaves
        self.tree_depth = tree_depth
        self.float_type = float_type
        self.splitting_type = splitting_type
        self.n_jobs = n_jobs
        self.verbose = verbose
        self.device = device
        self._handle = None

    def __del__(self):
        try:
            if self._handle is not None:
                check("annf_free_resources", lib().annf_free_resources(self._handle.h))
                self._handle.close()
        except Exception:
            pass

    def get_params(self):
        return {
            "n_neighbors": self.n_neighbors,
            "leaf_size": self.leaf_size,
            "max_leaves": self.max_leaves,
            "tree_depth": self.tree_depth,
            "float_type": self.float_type,
            "splitting_type": self.splitting_type,
            "n_jobs": self.n_jobs,
            "verbose": self.verbose,
        }

    # ------------------------------------------------------------------
    def _handle_for_device(self):
        if self._handle is None:
            self._handle = Handle(self.device)
        return self._handle

    def _check_float_type(self):
        assert self.float_type in self.ALLOWED_FLOAT_TYPES
        if self.float_type != "float":
            raise Exception("float_type '%s' is not available: the HIP backend computes in float32"
                            % self.float_type)
        self.numpy_dtype_float = np.float32
        self.numpy_dtype_int = np.int32

    def _init_native(self, final_tree_depth):
        if self.max_leaves is None:
            final_max_leaves = 2147483640
        else:
            if self.max_leaves < 0:
                raise Exception("max_leaves must be a positive integer")
            final_max_leaves = int(self.max_leaves)
        n_jobs = multiprocessing.cpu_count() if self.n_jobs < 1 else self.n_jobs
        h = self._handle_for_device()
        check("annf_init", lib().annf_init(h.h, int(self.n_neighbors), int(final_tree_depth), int(final_max_leaves),
                                           int(n_jobs), self.SPLITTING_TYPE_MAPPINGS[self.splitting_type],
                                           int(self.verbose > 0)))
        return h

    def fit(self, X):
        """kdtree/base.py:92-153."""
        self._check_float_type()
        final_tree_depth = self._compute_final_tree_depth(len(X), self.leaf_size, self.tree_depth)
        if self.verbose > 0:
            print("Final tree depth: " + str(final_tree_depth))
        X = np.ascontiguousarray(X)
        self.X = np.ascontiguousarray(X.astype(self.numpy_dtype_float))
        h = self._init_native(final_tree_depth)
        n, d = self.X.shape
        check("annf_fit", lib().annf_fit(h.h, ptr(self.X, _lib._fp), n, d))
        self._final_tree_depth = final_tree_depth
        return self

    def kneighbors(self, X=None, n_neighbors=None, return_distance=True):
        """kdtree/base.py:155-208: exact k-NN; returns (sqrt distances, original indices)."""
        assert return_distance == True  # noqa: E712 (reference behaviour)
        if n_neighbors is not None and self.n_neighbors != n_neighbors:
            self.n_neighbors = n_neighbors
            if self.verbose > 0:
                print("\nMODEL MUST BE RETRAINED ...\n")
            self.fit(self.X)
        if X is None:
            X = self.X
        else:
            X = np.ascontiguousarray(np.ascontiguousarray(X).astype(self.numpy_dtype_float))
        K = self.n_neighbors
        d_mins = np.zeros((X.shape[0], K), dtype=self.numpy_dtype_float)
        idx_mins = np.zeros((X.shape[0], K), dtype=self.numpy_dtype_int)
        check("annf_neighbors", lib().annf_neighbors(
            self._handle.h, ptr(X, _lib._fp), X.shape[0], X.shape[1], ptr(d_mins, _lib._fp), X.shape[0], K,
            ptr(idx_mins, _lib._ip), X.shape[0], K))
        return d_mins, idx_mins

    def process_rows(self, nn_indices, patches_a, patches_b, patchesred, n_rows, n_cols, n_neighbors=None,
                     propagate=True, select_best_nn=True):
        """kdtree/base.py:210-241: fills nn_indices (int32, in place)."""
        patches_a = np.ascontiguousarray(np.ascontiguousarray(patches_a).astype(np.uint8))
        patches_b = np.ascontiguousarray(np.ascontiguousarray(patches_b).astype(np.uint8))
        patchesred = np.ascontiguousarray(np.ascontiguousarray(patchesred).astype(self.numpy_dtype_float))
        if n_neighbors is not None and self.n_neighbors != n_neighbors:
            self.n_neighbors = n_neighbors
            if self.verbose > 0:
                print("\nMODEL MUST BE RETRAINED ...\n")
            self.fit(self.X)
        if not (isinstance(nn_indices, np.ndarray) and nn_indices.dtype == np.int32 and nn_indices.flags.c_contiguous):
            raise TypeError("nn_indices must be a C-contiguous int32 array (updated in place)")
        K = self.n_neighbors if n_neighbors is None else n_neighbors
        check("annf_process_rows", lib().annf_process_rows(
            self._handle.h, ptr(nn_indices, _lib._ip), nn_indices.shape[0],
            ptr(patches_a, _lib._u8p), patches_a.shape[0], patches_a.shape[1],
            ptr(patches_b, _lib._u8p), patches_b.shape[0], patches_b.shape[1],
            ptr(patchesred, _lib._fp), patchesred.shape[0], patchesred.shape[1],
            int(n_rows), int(n_cols), int(K), int(propagate), int(select_best_nn)))

    # ------------------------------------------------------------------
    def tree_arrays(self):
        """The fitted tree: axis/split per internal node, leaves ([from, to)
        per leaf), xi_orig/xi_pts (tree-position order) and pli (leaf of each
        original index) -- the arrays of the reference's KD_TREE_RECORD."""
        h = self._handle
        n, d, depth = C_int3(h)
        nn = (1 << depth) - 1
        out = {
            "axis": np.zeros(nn, np.int32), "split": np.zeros(nn, np.float32),
            "leaves": np.zeros(2 << depth, np.int32), "xi_orig": np.zeros(n, np.int32),
            "xi_pts": np.zeros((n, d), np.float32), "pli": np.zeros(n, np.int32),
        }
        check("annf_tree_export", lib().annf_tree_export(
            h.h, ptr(out["axis"], _lib._ip), ptr(out["split"], _lib._fp), ptr(out["leaves"], _lib._ip),
            ptr(out["xi_orig"], _lib._ip), ptr(out["xi_pts"], _lib._fp), ptr(out["pli"], _lib._ip)))
        out["depth"] = depth
        return out

    def save(self, path):
        """Writes the fitted tree and the constructor parameters to ``path``
        (the ``annfmp_amd.kdtree`` version-1 ``.npz`` of ``formats``)."""
        from . import formats
        params = self.get_params()
        params["device"] = self.device
        formats.save_tree(path, self.tree_arrays(), params)

    @classmethod
    def load(cls, path, device=None):
        """A KDTree restored from ``save``'s file, ready for ``kneighbors``
        without refitting; ``X`` is rebuilt from the points in tree order."""
        from . import formats
        arr = formats.load_tree(path)
        p = dict(arr["params"])
        if device is not None:
            p["device"] = device
        known = ("n_neighbors", "leaf_size", "max_leaves", "tree_depth", "float_type", "splitting_type",
                 "n_jobs", "verbose", "device")
        t = cls(**{k: p[k] for k in known if k in p})
        t._check_float_type()
        h = t._init_native(arr["depth"])
        check("annf_tree_import", lib().annf_tree_import(
            h.h, arr["n"], arr["d"], arr["depth"], ptr(arr["axis"], _lib._ip), ptr(arr["split"], _lib._fp),
            ptr(arr["leaves"], _lib._ip), ptr(arr["xi_orig"], _lib._ip), ptr(arr["xi_pts"], _lib._fp),
            ptr(arr["pli"], _lib._ip)))
        X = np.empty_like(arr["xi_pts"])
        X[arr["xi_orig"]] = arr["xi_pts"]
        t.X = X
        t._final_tree_depth = arr["depth"]
        return t

    def _compute_final_tree_depth(self, n, leaf_size, tree_depth):
        """kdtree/base.py:260-300 (same float log semantics)."""
        d = int(math.floor(math.log(n, 2)))
        if tree_depth is not None:
            if tree_depth > d:
                warnings.warn("tree_depth %i too large (using smaller depth): %i.\n" % (tree_depth, d))
                return d
            return tree_depth
        if not self.leaf_size > 1:
            raise Exception("Parameter 'leaf_size' must be larger than 1")
        subtr = int(math.floor(math.log(leaf_size, 2)))
        if d - subtr < 0:
            warnings.warn("tree_depth smaller than 0; setting tree_depth to 0")
            return 0
        return d - subtr

    def __repr__(self):
        return str(self.get_params())


def C_int3(h):
    import ctypes as C
    n, d, depth = C.c_int(), C.c_int(), C.c_int()
    check("annf_tree_info", lib().annf_tree_info(h.h, C.byref(n), C.byref(d), C.byref(depth)))
    return n.value, d.value, depth.value


def tree_depth_for(n, leaf_size, tree_depth=None):
    """Module-level form of KDTree._compute_final_tree_depth."""
    return KDTree(leaf_size=leaf_size)._compute_final_tree_depth(n, leaf_size, tree_depth)
