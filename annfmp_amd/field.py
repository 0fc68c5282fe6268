"""ANNField-compatible front (annfmp/base.py:13-142).

Same constructor, ``get_params`` / ``set_params``, ``fit(image_a, image_b)``
and ``patches_a`` / ``patches_b``; extra keyword arguments are passed to the
backend (``leaf_distance``, ``encoder``, ``shortlist``, ``nprobe``,
``projection``, ``device``).  The only algorithm provided here is
``"hip"`` (the MI355X backend, ``ANNFieldPropKDTreeHIP``); the reference's
other algorithm names raise, as an unknown name does there (base.py:72-73).
"""
from __future__ import annotations

from .base import ANNFieldPropKDTreeHIP

ALGORITHMS = ["hip"]
REFERENCE_ONLY = ["sklearn", "naive", "openmp", "openclopt"]


class ANNField:
    def __init__(self, n_neighbors=1, psize=8, dim_reduced=24, n_subset=1000, algorithm="hip", leaf_size=64,
                 n_jobs=-1, propagate=True, select_best_nn=True, verbose=1, seed=0, **backend_options):
        self.n_neighbors = n_neighbors
        self.psize = psize
        self.dim_reduced = dim_reduced
        self.n_subset = n_subset
        self.algorithm = algorithm
        self.leaf_size = leaf_size
        self.n_jobs = n_jobs
        self.propagate = propagate
        self.select_best_nn = select_best_nn
        self.verbose = verbose
        self.seed = seed
        # options of ANNFieldPropKDTreeHIP beyond the reference's (leaf_distance, encoder, ...)
        self.backend_options = dict(backend_options)

    def get_params(self, deep=True):
        return {
            "n_neighbors": self.n_neighbors,
            "psize": self.psize,
            "dim_reduced": self.dim_reduced,
            "n_subset": self.n_subset,
            "algorithm": self.algorithm,
            "leaf_size": self.leaf_size,
            "n_jobs": self.n_jobs,
            "propagate": self.propagate,
            "select_best_nn": self.select_best_nn,
            "verbose": self.verbose,
            "seed": self.seed,
        }

    def set_params(self, **parameters):
        for parameter, value in parameters.items():
            setattr(self, parameter, value)
        return self

    def fit(self, image_a, image_b):
        if self.algorithm in REFERENCE_ONLY:
            raise Exception("Algorithm '{}' belongs to the reference package; this package provides {}"
                            .format(self.algorithm, ALGORITHMS))
        if self.algorithm not in ALGORITHMS:
            raise Exception("Unknown algorithm '{}'!".format(self.algorithm))
        self._model = ANNFieldPropKDTreeHIP(
            n_neighbors=self.n_neighbors, psize=self.psize, dim_reduced=self.dim_reduced, n_subset=self.n_subset,
            leaf_size=self.leaf_size, n_jobs=self.n_jobs, propagate=self.propagate,
            select_best_nn=self.select_best_nn, verbose=self.verbose, seed=self.seed, **self.backend_options)
        return self._model.fit(image_a, image_b)

    @property
    def patches_a(self):
        return self._model.patches_a

    @property
    def patches_b(self):
        return self._model.patches_b
