"""GPU test of annf_mlp_triplet_grad_noalloc_dev, the entry point that
allocates nothing (its workspace is a fixed per-GPU device array): called
through the C ABI, its gradients and loss must equal those of
annf_mlp_triplet_grad_dev bit for bit (the same kernels on the same inputs),
on two calls in a row, the second with a smaller batch (the array keeps the
first call's contents), and a batch beyond the array is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(rng, B, D):
    q = rng.integers(0, 256, (B, D))
    p = np.clip(q + rng.integers(-20, 21, (B, D)), 0, 255)
    n = rng.integers(0, 256, (B, D))
    return np.concatenate([q, p, n]).astype(np.uint8)


def test_noalloc_triplet_grad_equals_the_allocating_entry(gpu_available):
    import torch
    from annfmp_amd import _lib
    from annfmp_amd import device as dev
    from oracle import mlp_oracle as M
    L = _lib.lib()
    D, h1, h2, d = 363, 128, 128, 16
    params = M.init_params(D, h1, h2, d, seed=3).astype(np.float32)
    n = params.size
    tp = torch.from_numpy(params).to("cuda")
    rng = np.random.default_rng(21)
    h = _lib.Handle(0)
    for B in (1024, 300):
        x = torch.from_numpy(_batch(rng, B, D)).to("cuda")
        g_ref = torch.zeros(n, dtype=torch.float32, device="cuda")
        l_ref = torch.zeros(1, dtype=torch.float32, device="cuda")
        g_new = torch.full((n,), 7.0, dtype=torch.float32, device="cuda")
        l_new = torch.full((1,), 7.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        print(f"noalloc triplet grad: B={B}", flush=True)
        _lib.check("ref", L.annf_mlp_triplet_grad_dev(h.h, dev.dptr(x), B, D, dev.dptr(tp), h1, h2, d, 0.1,
                                                      dev.dptr(g_ref), dev.dptr(l_ref)))
        _lib.check("noalloc", L.annf_mlp_triplet_grad_noalloc_dev(h.h, dev.dptr(x), B, D, dev.dptr(tp), h1, h2, d,
                                                                  0.1, dev.dptr(g_new), dev.dptr(l_new)))
        a, b = g_ref.cpu().numpy(), g_new.cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (B, int((a != b).sum()))
        assert np.array_equal(l_ref.cpu().numpy().view(np.uint32), l_new.cpu().numpy().view(np.uint32))
        assert np.abs(a).max() > 0


def test_noalloc_triplet_grad_refuses_batches_beyond_its_array(gpu_available):
    from annfmp_amd import _lib
    L = _lib.lib()
    h = _lib.Handle(0)
    # 3 * 4096 rows of a 128-128-16 network need 20 M floats > 4 Mi; refused
    # before any pointer is used
    rc = L.annf_mlp_triplet_grad_noalloc_dev(h.h, 16, 4096, 363, 16, 128, 128, 16, 0.1, 16, None)
    assert rc == -5   # ANNF_EUNSUPPORTED
    assert b"allocation-free" in L.annf_last_error()
