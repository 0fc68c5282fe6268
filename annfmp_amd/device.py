"""Device-memory plumbing: torch tensors as HBM buffers for the C ABI.

PyTorch only allocates and moves data here; every computation on those
buffers is a kernel of libannf_hip.so.
"""
from __future__ import annotations

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def cuda_device(index=0):
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("no GPU visible: the HIP backend needs an MI355X (gfx950) device")
    return t.device("cuda", index)


def dptr(tensor):
    """Raw device pointer of a tensor (None -> NULL)."""
    if tensor is None:
        return None
    if not tensor.is_contiguous():
        raise ValueError("device buffers passed to libannf_hip.so must be contiguous")
    return tensor.data_ptr()
