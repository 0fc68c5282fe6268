"""ctypes binding of libannf_hip.so (C ABI in include/annf_hip.h).

The library is required: there is no CPU fallback.  Loading fails loudly
when the shared object is missing (build it with ``python -m
annfmp_amd.build``) and every entry point raises ``ANNFError`` on a
non-zero status.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libannf_hip.so")

ANNF_OK = 0
ERRORS = {-1: "EINVAL", -2: "ENOTREE", -3: "EHIP", -4: "ENOMEM", -5: "EUNSUPPORTED"}
MAX_K = 64
MAX_D = 64

# every symbol include/annf_hip.h declares
EXPORTS = (
    "annf_create", "annf_destroy", "annf_last_error", "annf_stream", "annf_init", "annf_fit", "annf_fit_dev",
    "annf_neighbors", "annf_process_rows", "annf_norm_array", "annf_free_resources", "annf_tree_info",
    "annf_tree_export", "annf_tree_import", "annf_field_dev", "annf_stage_times", "annf_set_timing",
    "annf_exact_field_dev",
    "annf_field_mm_dev", "annf_mlp_forward_dev", "annf_mlp_triplet_grad_dev", "annf_mlp_adam_dev",
    "annf_set_encoder", "annf_encode_image_dev", "annf_pca_cov_dev", "annf_mlp_triplet_grad_noalloc_dev",
    "annf_pca_fit_dev", "annf_patch_fp8_dev", "annf_project_dev", "annf_field_openclopt_dev", "annf_openclopt_tree_export",
)

PROJ_PCA32 = 0
PROJ_PCA64 = 1
PROJ_ENCODER = 2

PREC_BF16 = 0
PREC_FP8 = 1


class ANNFError(RuntimeError):
    """A libannf_hip.so call returned an error status."""

    def __init__(self, func, code, msg):
        self.func, self.code = func, code
        super().__init__(f"{func}: {ERRORS.get(code, code)}: {msg}")


class HipLibraryMissing(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()

_vp = C.c_void_p
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


def _declare(L):
    i = C.c_int
    L.annf_create.argtypes = [C.POINTER(_vp), i]
    L.annf_destroy.argtypes = [_vp]
    L.annf_last_error.restype = C.c_char_p
    L.annf_last_error.argtypes = []
    L.annf_stream.restype = _vp
    L.annf_stream.argtypes = [_vp]
    L.annf_init.argtypes = [_vp, i, i, i, i, i, i]
    L.annf_fit.argtypes = [_vp, _fp, i, i]
    L.annf_fit_dev.argtypes = [_vp, _vp, i, i]
    L.annf_neighbors.argtypes = [_vp, _fp, i, i, _fp, i, i, _ip, i, i]
    L.annf_process_rows.argtypes = [_vp, _ip, i, _u8p, i, i, _u8p, i, i, _fp, i, i, i, i, i, i, i]
    L.annf_norm_array.restype = C.c_float
    L.annf_norm_array.argtypes = [_u8p, _u8p, i]
    L.annf_free_resources.argtypes = [_vp]
    L.annf_tree_info.argtypes = [_vp, _ip, _ip, _ip]
    L.annf_tree_export.argtypes = [_vp, _ip, _fp, _ip, _ip, _fp, _ip]
    L.annf_tree_import.argtypes = [_vp, i, i, i, _ip, _fp, _ip, _ip, _fp, _ip]
    L.annf_field_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, _vp, _vp, i, i, i, i, i, i, _vp, _vp, _vp, _vp, _vp]
    L.annf_stage_times.argtypes = [_vp, _fp, i]
    L.annf_set_timing.argtypes = [_vp, i]
    L.annf_exact_field_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, _vp, i, i, _vp, _vp]
    L.annf_field_mm_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, _vp, _vp, i, i, i, i, i, i, i, i, _vp, _vp,
                                    _vp, _vp, _vp]
    f = C.c_float
    L.annf_mlp_forward_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, _vp]
    L.annf_mlp_triplet_grad_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, f, _vp, _vp]
    L.annf_mlp_triplet_grad_noalloc_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, f, _vp, _vp]
    L.annf_mlp_adam_dev.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_longlong, f, f, f, f, i]
    L.annf_set_encoder.argtypes = [_vp, _vp, i, i, i, i]
    L.annf_encode_image_dev.argtypes = [_vp, _vp, i, i, i, _vp]
    L.annf_pca_cov_dev.argtypes = [_vp, _vp, i, i, _vp, i, _vp, i, i, _vp, i, i, _vp, _vp]
    L.annf_patch_fp8_dev.argtypes = [_vp, _vp, i, i, i, _vp, _vp, _vp]
    L.annf_project_dev.argtypes = [_vp, _vp, i, i, i, _vp, _vp, i, i, _vp]
    L.annf_field_openclopt_dev.argtypes = [_vp, _vp, i, i, _vp, i, i, i, _vp, i, i, _vp, i, i, i, _vp, _vp, _vp]
    L.annf_openclopt_tree_export.argtypes = [_vp, _ip, _ip, _fp, _ip, _ip, _ip]
    L.annf_pca_fit_dev.argtypes = [_vp, _vp, i, i, _vp, i, _vp, i, i, _vp, i, i, i, _vp, _vp, _vp]
    for name in EXPORTS:
        if name not in ("annf_last_error", "annf_stream", "annf_norm_array"):
            getattr(L, name).restype = C.c_int


def lib(path=None):
    """Loads (once) and returns the library; raises HipLibraryMissing."""
    global _lib
    with _lock:
        if _lib is None:
            p = path or LIB_PATH
            if not os.path.exists(p):
                raise HipLibraryMissing(
                    f"{p} is missing: build it with `python -m annfmp_amd.build` (no CPU fallback exists)")
            # let torch bring up the device first: a HIP runtime started by
            # this library alone leaves torch without a visible GPU
            try:
                import torch
                if torch.cuda.is_available():
                    torch.cuda.init()
            except ImportError:
                pass
            L = C.CDLL(p)
            _declare(L)
            _lib = L
        return _lib


def check(func, rc):
    if rc != ANNF_OK:
        msg = lib().annf_last_error()
        raise ANNFError(func, rc, msg.decode() if msg else "")
    return rc


def ptr(a, t):
    return a.ctypes.data_as(t)


class Handle:
    """An annf_ctx on one device."""

    def __init__(self, device=0):
        self._h = None
        h = _vp()
        check("annf_create", lib().annf_create(C.byref(h), int(device)))
        self._h = h
        self.device = device

    @property
    def h(self):
        if self._h is None:
            raise ANNFError("handle", -1, "handle already destroyed")
        return self._h

    def stream(self):
        return lib().annf_stream(self.h)

    def close(self):
        if self._h is not None:
            lib().annf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
