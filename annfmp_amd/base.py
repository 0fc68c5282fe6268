"""ANNFieldPropKDTreeHIP: the openmp backend's interface on MI355X.

Mirrors ``ANNFieldPropKDTreeOpenMP`` (annfmp/openmp/base.py:16-219): same
constructor parameters, ``fit(image_a, image_b) -> nn_indices`` (int32,
``(N,)``, ``nn_indices[y * n_cols + x]`` = original index of the matched B
patch), ``patches_a`` / ``patches_b`` and the ``_timers`` keys.

The PCA basis is fitted exactly as the reference does (scikit-learn
``extract_patches_2d`` samples of ``n_subset // 2`` patches from A with
``random_state=seed`` and from B with ``seed + 1``, then ``PCA.fit``,
openmp/base.py:168-191).  Everything after the fit runs on the GPU through
``annf_field_dev``: patch extraction fused with the projection, the tree over
B, row 0, the propagation sweep and best-of-K.  Patches are never
materialised on the host unless ``patches_a`` / ``patches_b`` is read.

Extra keyword arguments (not in the reference): ``device`` (GPU ordinal),
``projection`` ("fp32": FP32 accumulation; "fp64": FP64 accumulation rounded
once to float32, which reproduces the reference's FP64 transform up to its
last-bit rounding), ``leaf_distance`` ("reduced": the reference's method,
distances in the projected space, the default; "bf16" / "fp8": the
full-patch leaf k-NN on MFMA of ``annf_field_mm_dev``, its list ranked by the
full-patch distance, propagation as the parallel second pass), ``encoder``
(a ``PatchEncoder`` or the path of its ``.npz``: the learned projection
replaces PCA), ``shortlist`` (FP8 re-rank length), ``nprobe`` (leaves per
query for "bf16" / "fp8") and ``encoder_check`` (with an encoder: when > 0,
the field is also computed with the PCA projection, recall@K of both fields is
measured on that many sampled A patches against the exact field, and the
better one is returned -- PCA is the fallback when the encoder's recall is
lower), ``pca_fit`` ("sklearn": scikit-learn's ``PCA.fit`` as the
reference, the default; "gpu": the same sampled patches, FP64 covariance and
eigensolve on the GPU through ``annf_pca_fit_dev`` -- scikit-learn's
covariance_eigh solution whatever solver sklearn would pick) and
``semantics`` ("openmp": the openmp backend's tree and rows, the default;
"openclopt": the openclopt backend's, annfmp/openclopt/base.py:47-130 --
padded power-of-two tree, K fixed to 8, propagation and best-of-K always on,
see ``annfmp_amd.openclopt``).  Extra outputs after ``fit``: ``nn_distances`` (full-patch L2 of each
chosen match, float32), ``knn_indices`` (the K candidates of every patch before
best-of-K, original indices), ``projection_used`` ("pca" or "encoder") and,
with ``encoder_check``, ``encoder_check_result`` (both recalls).
"""
from __future__ import annotations

import collections
import time

import numpy

from . import device as dev
from . import _lib
from .kdtree import tree_depth_for


class ANNFieldPropKDTreeHIP:
    def __init__(self, n_neighbors=3, psize=8, dim_reduced=5, n_subset=1000, leaf_size=64, n_jobs=-1,
                 propagate=True, select_best_nn=True, float_type="float", verbose=1, seed=0, device=0,
                 projection="fp32", leaf_distance="reduced", encoder=None, shortlist=20, nprobe=1,
                 encoder_check=0, pca_fit="sklearn", semantics="openmp"):
        self.n_neighbors = n_neighbors
        self.psize = psize
        self.dim_reduced = dim_reduced
        self.n_subset = n_subset
        self.leaf_size = leaf_size
        self.n_jobs = n_jobs
        self.propagate = propagate
        self.select_best_nn = select_best_nn
        self.float_type = float_type
        self.verbose = verbose
        self.seed = seed
        self.device = device
        self.projection = projection
        if leaf_distance not in ("reduced", "bf16", "fp8"):
            raise Exception("leaf_distance must be 'reduced', 'bf16' or 'fp8', not {!r}".format(leaf_distance))
        self.leaf_distance = leaf_distance
        if isinstance(encoder, str):
            from .encoder import PatchEncoder
            encoder = PatchEncoder.load(encoder)
        if encoder is not None and encoder.psize != psize:
            raise Exception("encoder trained for patch size {}, psize is {}".format(encoder.psize, psize))
        self.encoder = encoder
        self.shortlist = shortlist
        self.nprobe = nprobe
        self.encoder_check = int(encoder_check)
        if pca_fit not in ("sklearn", "gpu"):
            raise Exception("pca_fit must be 'sklearn' or 'gpu', not {!r}".format(pca_fit))
        self.pca_fit = pca_fit
        if semantics not in ("openmp", "openclopt"):
            raise Exception("semantics must be 'openmp' or 'openclopt', not {!r}".format(semantics))
        self.semantics = semantics
        self._handle = None

    # ------------------------------------------------------------------
    def fit(self, image_a, image_b):
        numpy.random.seed(self.seed)
        self._get_numpy_float_type()
        self.image_a = image_a
        self.image_b = image_b
        self._timers = collections.OrderedDict()
        self._n_cols = self.image_a.shape[1] - self.psize + 1
        self._n_rows = self.image_a.shape[0] - self.psize + 1

        check = self.encoder is not None and self.encoder_check > 0
        tstart = time.time()
        need_pca = self.encoder is None or check
        self._pca_model = self._fit_pca_model(self.image_a, self.image_b) if need_pca else None
        self._timers["pca_fit"] = time.time() - tstart

        if self.semantics == "openclopt":
            return self._fit_openclopt(image_a, image_b)

        # since there is just one leaf visit, this must hold (openmp/base.py:113)
        assert self.leaf_size > self.n_neighbors

        n_b = (image_b.shape[0] - self.psize + 1) * (image_b.shape[1] - self.psize + 1)
        depth = tree_depth_for(n_b, self.leaf_size)
        comp = self._pca_model.components_ if self._pca_model is not None else None
        mean = self._pca_model.mean_ if self._pca_model is not None else None
        res = self._field(image_a, image_b, comp, mean, depth,
                          self.encoder if self.encoder is not None else self.projection)
        self.projection_used = "encoder" if self.encoder is not None else "pca"
        self.encoder_check_result = None
        if check:
            # the fallback: PCA's field on the same pair, and recall@K of both
            # against the exact field on sampled A patches
            from . import metrics
            res_pca = self._field(image_a, image_b, comp, mean, depth, self.projection)
            k = self.n_neighbors
            q = metrics.sample_queries(res["nn"].shape[0], self.encoder_check, seed=self.seed)
            exact = metrics.exact_knn(image_a, image_b, self.psize, q, k, device=self.device,
                                      handle=self._get_handle())
            r_enc = metrics.field_quality(image_a, image_b, self.psize, res["nn"], res["knn"], k=k,
                                          n_sample=q.size, seed=self.seed, exact=exact)["recall@%d" % k]
            r_pca = metrics.field_quality(image_a, image_b, self.psize, res_pca["nn"], res_pca["knn"], k=k,
                                          n_sample=q.size, seed=self.seed, exact=exact)["recall@%d" % k]
            self.encoder_check_result = {"encoder": r_enc, "pca": r_pca, "sample": int(q.size)}
            if r_enc < r_pca:
                res = res_pca
                self.projection_used = "pca"
        nn_indices = res["nn"]
        self.nn_distances = res["dist"]
        self.knn_indices = res["knn"]
        st = res["stages_ms"]
        self._timers["extracting_patches"] = 0.0   # fused into the projection
        # A is projected on a second stream during the tree build, so its
        # projection time is inside kdtree_fit
        self._timers["pca_apply"] = (st[0] + st[1]) / 1000.0
        self._timers["kdtree_fit"] = st[2] / 1000.0
        self._timers["remaining_rows"] = sum(st[3:]) / 1000.0
        self._timers["device_total"] = res["wall_s"]
        if self.verbose > 0:
            print("-----------------------------------------")
            print("--------------- RUNTIMES ----------------")
            print("-----------------------------------------")
            print("Extraction of patches:\t{}".format(self._timers["extracting_patches"]))
            print("PCA fit:\t\t{}".format(self._timers["pca_fit"]))
            print("PCA apply:\t\t{}".format(self._timers["pca_apply"]))
            print("K-D Tree (fitting):\t{}".format(self._timers["kdtree_fit"]))
            print("Remaining rows:\t\t{}".format(self._timers["remaining_rows"]))
            print("Total runtime:\t\t{}".format(sum(self._timers[k] for k in
                                                     ("extracting_patches", "pca_fit", "pca_apply",
                                                      "kdtree_fit", "remaining_rows"))))
            print("-----------------------------------------")
        return nn_indices

    def _fit_openclopt(self, image_a, image_b):
        """openclopt/base.py:47-130 after the PCA fit: the reference asserts
        select_best_nn and fixes K = 8 whatever n_neighbors says."""
        from .openclopt import field_openclopt
        assert self.select_best_nn
        if self.encoder is not None:
            raise Exception("semantics='openclopt' projects with PCA (the reference's reducePatchDim)")
        tstart = time.time()
        res = field_openclopt(image_a, image_b, self.psize, self._pca_model.components_, self._pca_model.mean_,
                              self.leaf_size, projection=self.projection, handle=self._get_handle(),
                              device=self.device)
        self._timers["futhark"] = time.time() - tstart
        self.projection_used = "pca"
        self.encoder_check_result = None
        self.knn_indices = res["knn"]
        self.nn_distances = None
        if self.verbose > 0:
            print("-----------------------------------------")
            print("--------------- RUNTIMES ----------------")
            print("-----------------------------------------")
            print("PCA fit:\t\t{}".format(self._timers["pca_fit"]))
            print("Futhark:\t\t{}".format(self._timers["futhark"]))
            print("Total runtime:\t\t{}".format(self._timers["pca_fit"] + self._timers["futhark"]))
            print("-----------------------------------------")
        return res["nn"]

    def _field(self, image_a, image_b, comp, mean, depth, proj):
        if self.leaf_distance == "reduced":
            return field_on_device(image_a, image_b, self.psize, comp, mean, depth, self.n_neighbors,
                                   self.propagate, self.select_best_nn, projection=proj, device=self.device,
                                   handle=self._get_handle())
        return field_mm_on_device(image_a, image_b, self.psize, comp, mean, depth, self.n_neighbors,
                                  nprobe=self.nprobe, precision=self.leaf_distance, shortlist=self.shortlist,
                                  projection=proj, propagate=bool(self.propagate), device=self.device,
                                  handle=self._get_handle())

    def _get_handle(self):
        if self._handle is None:
            self._handle = _lib.Handle(self.device)
        return self._handle

    def _create_patches(self, img):
        from sklearn.feature_extraction import image
        n_channels = img.shape[-1]
        patches = image.extract_patches_2d(img, (self.psize, self.psize))
        return patches.reshape((-1, self.psize * self.psize * n_channels))

    def _fit_pca_model(self, image_a, image_b):
        """openmp/base.py:168-191 (scikit-learn, FP64); with pca_fit="gpu" the
        same samples fitted by annf_pca_fit_dev."""
        if self.pca_fit == "gpu":
            from .pca import fit_pca_gpu, sklearn_sample_indices
            n = int(self.n_subset / 2)
            p = self.psize
            ia = sklearn_sample_indices(image_a.shape[0], image_a.shape[1], p, n, self.seed)
            ib = sklearn_sample_indices(image_b.shape[0], image_b.shape[1], p, n, self.seed + 1)
            comp, mean, var = fit_pca_gpu(image_a, ia, p, self.dim_reduced, image_b, ib, device=self.device,
                                          handle=self._get_handle(), return_variance=True)
            return _GpuPCA(comp, mean, var)
        from sklearn.feature_extraction import image
        from sklearn.decomposition import PCA
        subset_a = image.extract_patches_2d(image_a, (self.psize, self.psize), max_patches=int(self.n_subset / 2),
                                            random_state=self.seed)
        subset_b = image.extract_patches_2d(image_b, (self.psize, self.psize), max_patches=int(self.n_subset / 2),
                                            random_state=self.seed + 1)
        subset_a = subset_a.reshape((subset_a.shape[0], -1))
        subset_b = subset_b.reshape((subset_b.shape[0], -1))
        both_subsets = numpy.concatenate((subset_a, subset_b), axis=0)
        model = PCA(n_components=self.dim_reduced)
        model.fit(both_subsets)
        return model

    def _get_numpy_float_type(self):
        if self.float_type == "float":
            return numpy.float32
        elif self.float_type == "double":
            raise Exception("float_type 'double' is not available: the HIP backend computes in float32")
        raise Exception("Unknown float type: {}".format(self.float_type))

    @property
    def patches_a(self):
        return self._create_patches(self.image_a)

    @property
    def patches_b(self):
        return self._create_patches(self.image_b)


class _GpuPCA:
    """The attributes of a fitted sklearn PCA the backend reads."""

    def __init__(self, components, mean, variance):
        self.components_ = components
        self.mean_ = mean
        self.explained_variance_ = variance
        self.n_components_ = components.shape[0]


def _projection_args(projection, components, mean, psize, handle, device):
    """projection "fp32" / "fp64" (PCA components + mean) or a PatchEncoder,
    installed on the handle.  Returns (basis, mean, d, ANNF_PROJ_*)."""
    torch = dev.torch()
    cuda = dev.cuda_device(device)
    if not isinstance(projection, str):
        enc = projection
        if enc.psize != psize:
            raise ValueError(f"encoder for patch size {enc.psize}, field uses {psize}")
        enc.install(handle, device)
        return None, None, enc.d, _lib.PROJ_ENCODER
    if projection not in ("fp32", "fp64"):
        raise ValueError(f"projection {projection!r}: 'fp32', 'fp64' or a PatchEncoder")
    # the FP64 projection takes the basis as given (float64), the FP32 one rounded
    ftype = numpy.float64 if projection == "fp64" else numpy.float32
    comp = torch.from_numpy(numpy.ascontiguousarray(components, dtype=ftype)).to(cuda)
    mu = torch.from_numpy(numpy.ascontiguousarray(mean, dtype=numpy.float64)).to(cuda)
    return comp, mu, comp.shape[0], _lib.PROJ_PCA64 if projection == "fp64" else _lib.PROJ_PCA32


def field_on_device(image_a, image_b, psize, components, mean, tree_depth, n_neighbors, propagate=True,
                    select_best_nn=True, projection="fp32", device=0, handle=None, return_reduced=False,
                    timing=True):
    """One call of annf_field_dev on host images; returns numpy results.

    components: (d, p*p*3) basis, mean: (p*p*3,) -- sklearn PCA layout."""
    torch = dev.torch()
    h = handle or _lib.Handle(device)
    L = _lib.lib()
    cuda = dev.cuda_device(device)
    ia = torch.from_numpy(numpy.ascontiguousarray(image_a, dtype=numpy.uint8)).to(cuda)
    ib = torch.from_numpy(numpy.ascontiguousarray(image_b, dtype=numpy.uint8)).to(cuda)
    comp, mu, d, proj = _projection_args(projection, components, mean, psize, h, device)
    ha, wa = image_a.shape[:2]
    hb, wb = image_b.shape[:2]
    n_a = (ha - psize + 1) * (wa - psize + 1)
    n_b = (hb - psize + 1) * (wb - psize + 1)
    K = int(n_neighbors)
    nn = torch.empty(n_a, dtype=torch.int32, device=cuda)
    dist = torch.empty(n_a, dtype=torch.float32, device=cuda)
    knn = torch.empty((n_a, K), dtype=torch.int32, device=cuda)
    ra = torch.empty((n_a, d), dtype=torch.float32, device=cuda) if return_reduced else None
    rb = torch.empty((n_b, d), dtype=torch.float32, device=cuda) if return_reduced else None
    torch.cuda.synchronize(cuda)
    _lib.check("annf_set_timing", L.annf_set_timing(h.h, int(bool(timing))))
    t0 = time.time()
    _lib.check("annf_field_dev", L.annf_field_dev(
        h.h, dev.dptr(ia), ha, wa, dev.dptr(ib), hb, wb, int(psize), dev.dptr(comp), dev.dptr(mu), int(d),
        proj, int(tree_depth), K, int(bool(propagate)), int(bool(select_best_nn)),
        dev.dptr(nn), dev.dptr(dist), dev.dptr(knn), dev.dptr(ra), dev.dptr(rb)))
    wall = time.time() - t0
    st = numpy.zeros(7, numpy.float32)
    L.annf_stage_times(h.h, _lib.ptr(st, _lib._fp), 7)
    out = {"nn": nn.cpu().numpy(), "dist": dist.cpu().numpy(), "knn": knn.cpu().numpy(),
           "stages_ms": st.tolist(), "wall_s": wall}
    if return_reduced:
        out["red_a"] = ra.cpu().numpy()
        out["red_b"] = rb.cpu().numpy()
    return out


def fp8_patch_bytes(psize):
    """Bytes per patch of the FP8 codes (rows of ceil(3p / 4) dwords, padded
    to a multiple of 64)."""
    rw = (3 * psize + 3) // 4
    return (psize * rw * 4 + 63) // 64 * 64


def patch_fp8_on_device(image, psize, device=0, handle=None):
    """The FP8 leaf pass's operands of every patch of a host image
    (annf_patch_fp8_dev): (codes n x DQ uint8, norms, inverse scales)."""
    torch = dev.torch()
    h = handle or _lib.Handle(device)
    cuda = dev.cuda_device(device)
    H, W = image.shape[:2]
    n = (H - psize + 1) * (W - psize + 1)
    ti = torch.from_numpy(numpy.ascontiguousarray(image, dtype=numpy.uint8)).to(cuda)
    codes = torch.empty((n, fp8_patch_bytes(psize)), dtype=torch.uint8, device=cuda)
    nrm = torch.empty(n, dtype=torch.float32, device=cuda)
    inv = torch.empty(n, dtype=torch.float32, device=cuda)
    _lib.check("annf_patch_fp8_dev", _lib.lib().annf_patch_fp8_dev(
        h.h, dev.dptr(ti), H, W, int(psize), dev.dptr(codes), dev.dptr(nrm), dev.dptr(inv)))
    return codes.cpu().numpy(), nrm.cpu().numpy(), inv.cpu().numpy()


def field_mm_on_device(image_a, image_b, psize, components, mean, tree_depth, n_neighbors, nprobe=1,
                       precision="bf16", shortlist=20, projection="fp32", propagate=False, device=0, handle=None,
                       return_reduced=False, timing=True):
    """One call of annf_field_mm_dev (full-patch leaf k-NN on MFMA) on host
    images; returns numpy results like field_on_device.

    precision: "bf16" (exact distances) or "fp8" (per query, the shortlist
    smallest approximate distances over its scanned leaves, re-ranked
    exactly); nprobe: leaves per query (a power of two); propagate: False,
    True (all K entries of the list above propagate) or a count, from the
    first-pass lists of the row above (one parallel pass), or "chain" (all K,
    from the row above's final lists, rows in order: the reference's
    dependence; BF16 only)."""
    torch = dev.torch()
    h = handle or _lib.Handle(device)
    L = _lib.lib()
    cuda = dev.cuda_device(device)
    ia = torch.from_numpy(numpy.ascontiguousarray(image_a, dtype=numpy.uint8)).to(cuda)
    ib = torch.from_numpy(numpy.ascontiguousarray(image_b, dtype=numpy.uint8)).to(cuda)
    comp, mu, d, proj = _projection_args(projection, components, mean, psize, h, device)
    ha, wa = image_a.shape[:2]
    hb, wb = image_b.shape[:2]
    n_a = (ha - psize + 1) * (wa - psize + 1)
    n_b = (hb - psize + 1) * (wb - psize + 1)
    K = int(n_neighbors)
    prec = {"bf16": _lib.PREC_BF16, "fp8": _lib.PREC_FP8}[precision]
    nn = torch.empty(n_a, dtype=torch.int32, device=cuda)
    dist = torch.empty(n_a, dtype=torch.float32, device=cuda)
    knn = torch.empty((n_a, K), dtype=torch.int32, device=cuda)
    ra = torch.empty((n_a, d), dtype=torch.float32, device=cuda) if return_reduced else None
    rb = torch.empty((n_b, d), dtype=torch.float32, device=cuda) if return_reduced else None
    torch.cuda.synchronize(cuda)
    _lib.check("annf_set_timing", L.annf_set_timing(h.h, int(bool(timing))))
    t0 = time.time()
    _lib.check("annf_field_mm_dev", L.annf_field_mm_dev(
        h.h, dev.dptr(ia), ha, wa, dev.dptr(ib), hb, wb, int(psize), dev.dptr(comp), dev.dptr(mu), int(d),
        proj, int(tree_depth), K, int(nprobe), prec, int(shortlist),
        (K if propagate is True else (-K if propagate == "chain" else int(propagate or 0))), dev.dptr(nn),
        dev.dptr(dist), dev.dptr(knn), dev.dptr(ra), dev.dptr(rb)))
    wall = time.time() - t0
    st = numpy.zeros(7, numpy.float32)
    L.annf_stage_times(h.h, _lib.ptr(st, _lib._fp), 7)
    out = {"nn": nn.cpu().numpy(), "dist": dist.cpu().numpy(), "knn": knn.cpu().numpy(),
           "stages_ms": st[:6].tolist(), "wall_s": wall}
    if return_reduced:
        out["red_a"] = ra.cpu().numpy()
        out["red_b"] = rb.cpu().numpy()
    return out
