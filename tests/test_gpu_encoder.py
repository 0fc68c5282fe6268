"""GPU tests of the learned encoder kernels against the CPU oracle
(oracle/mlp_oracle.py, itself pinned by finite differences and torch
autograd): forward, triplet-loss gradients, Adam, the fused image encoder,
and fields whose tree is built on the encoder's embeddings.

Tolerances: the kernels round GEMM operands to BF16 and accumulate in FP32;
the oracle rounds at the same points but computes in FP64, so an activation
that lands next to a BF16 rounding boundary may round the other way.  The
bounds below (1e-2 relative in norm for gradients, 3e-3 of the output scale
for the forward pass) sit far below what a wrong term, sign or transpose
gives (order 1)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

D11, H1, H2, DOUT = 363, 128, 128, 16


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _handle():
    from annfmp_amd import _lib
    return _lib.Handle(0)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _triplet_batch(rng, B, D):
    q = rng.integers(0, 256, (B, D))
    p = np.clip(q + rng.integers(-20, 21, (B, D)), 0, 255)
    n = rng.integers(0, 256, (B, D))
    n[: B // 4] = np.clip(q[: B // 4] + rng.integers(-12, 13, (B // 4, D)), 0, 255)   # some hard negatives
    return np.concatenate([q, p, n]).astype(np.uint8)


def test_mlp_forward_matches_oracle(gpu_available):
    import torch
    from annfmp_amd import _lib
    from annfmp_amd import device as dev
    rng = np.random.default_rng(1)
    params = M.init_params(D11, H1, H2, DOUT, seed=1)
    x = rng.integers(0, 256, (700, D11)).astype(np.uint8)
    out = torch.empty((700, DOUT), dtype=torch.float32, device="cuda")
    h = _handle()
    tp, tx = _cuda(params), _cuda(x)
    _lib.check("fwd", _lib.lib().annf_mlp_forward_dev(h.h, dev.dptr(tx), 700, D11, dev.dptr(tp), H1, H2, DOUT,
                                                      dev.dptr(out)))
    ref = M.forward(x, params, D11, H1, H2, DOUT, bf16_ops=True)
    assert np.abs(out.cpu().numpy() - ref).max() <= 3e-3 * np.abs(ref).max()


@pytest.mark.parametrize("B,margin,seed", [(256, 0.1, 2), (100, 0.5, 3)])
def test_triplet_gradients_match_oracle(B, margin, seed, gpu_available):
    import torch
    from annfmp_amd import _lib
    from annfmp_amd import device as dev
    rng = np.random.default_rng(seed)
    params = M.init_params(D11, H1, H2, DOUT, seed=seed)
    x = _triplet_batch(rng, B, D11)
    g = torch.empty(params.size, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    h = _handle()
    tp, tx = _cuda(params), _cuda(x)
    _lib.check("grad", _lib.lib().annf_mlp_triplet_grad_dev(h.h, dev.dptr(tx), B, D11, dev.dptr(tp), H1, H2, DOUT,
                                                            margin, dev.dptr(g), dev.dptr(loss)))
    want_loss, want = M.triplet_grad(x, B, params, D11, H1, H2, DOUT, margin, bf16_ops=True)
    got = g.cpu().numpy()
    assert 0 < want_loss
    assert abs(loss.item() - want_loss) <= 1e-3 * want_loss
    o = 0
    for name, shape in zip(("W1", "b1", "W2", "b2", "W3"), [(H1, D11), (H1,), (H2, H1), (H2,), (DOUT, H2)]):
        n = int(np.prod(shape))
        err = _rel(got[o:o + n], want[o:o + n])
        print(f"{name}: relative gradient error {err:.2e}")
        assert err <= 1e-2, name
        o += n
    # the loss is invariant to a common shift of the embeddings, so the last
    # bias has no gradient: only FP32 rounding residue may remain
    gW3 = np.abs(want[o - DOUT * H2:o]).max()
    assert np.abs(want[o:]).max() <= 1e-12 * gW3
    assert np.abs(got[o:]).max() <= 1e-4 * gW3


def test_adam_matches_oracle(gpu_available):
    import torch
    from annfmp_amd import _lib
    from annfmp_amd import device as dev
    rng = np.random.default_rng(4)
    p = rng.normal(size=5000).astype(np.float32)
    tp = _cuda(p.copy())
    tm = torch.zeros(5000, dtype=torch.float32, device="cuda")
    tv = torch.zeros(5000, dtype=torch.float32, device="cuda")
    rp, rm, rv = p.astype(np.float64), np.zeros(5000), np.zeros(5000)
    h = _handle()
    for step in range(1, 4):
        g = rng.normal(size=5000).astype(np.float32)
        tg = _cuda(g)
        _lib.check("adam", _lib.lib().annf_mlp_adam_dev(h.h, dev.dptr(tp), dev.dptr(tg), dev.dptr(tm), dev.dptr(tv),
                                                        5000, 1e-3, 0.9, 0.999, 1e-8, step))
        rp, rm, rv = M.adam_step(rp, g.astype(np.float64), rm, rv, 1e-3, 0.9, 0.999, 1e-8, step)
    assert np.abs(tp.cpu().numpy() - rp).max() <= 1e-6


@pytest.mark.parametrize("p", [5, 7, 11])
def test_fused_encoder_matches_forward(p, gpu_available):
    """Fused extraction + MLP over an image against the unfused GPU forward
    on the patch rows and against the oracle."""
    import torch
    from annfmp_amd import _lib, synth
    from annfmp_amd import device as dev
    from annfmp_amd.encoder import PatchEncoder
    a, _ = synth.make_pair(synth.small_spec(57, 83), seed=p)
    enc = PatchEncoder(p, DOUT, seed=p)
    h = _handle()
    got = enc.encode_image(a, handle=h)
    pa = O.extract_patches(a, p)
    D = p * p * 3
    out = torch.empty((pa.shape[0], DOUT), dtype=torch.float32, device="cuda")
    tp, tx = _cuda(enc.params), _cuda(pa)
    _lib.check("fwd", _lib.lib().annf_mlp_forward_dev(h.h, dev.dptr(tx), pa.shape[0], D, dev.dptr(tp), H1, H2,
                                                      DOUT, dev.dptr(out)))
    unfused = out.cpu().numpy()
    ref = M.forward(pa, enc.params, D, H1, H2, DOUT, bf16_ops=True)
    scale = np.abs(ref).max()
    assert np.abs(got - unfused).max() <= 3e-3 * scale
    assert np.abs(got - ref).max() <= 3e-3 * scale


@pytest.mark.parametrize("d", [5, 3, 13])
def test_encode_image_row_stride_not_a_multiple_of_four(d, gpu_available):
    """annf_encode_image_dev writes n x d floats with row stride d: with
    d % 4 != 0 every row must equal the oracle's and nothing past the end of
    `out` may change (a guard region of sentinels follows it)."""
    import torch
    from annfmp_amd import _lib, synth
    from annfmp_amd import device as dev
    from annfmp_amd.encoder import PatchEncoder
    p = 5
    a, _ = synth.make_pair(synth.small_spec(41, 53), seed=d)
    enc = PatchEncoder(p, d, seed=d)
    h = _handle()
    enc.install(h)
    n = (a.shape[0] - p + 1) * (a.shape[1] - p + 1)
    guard = 64
    buf = torch.full((n * d + guard,), -7.25, dtype=torch.float32, device="cuda")
    img = _cuda(a)
    _lib.check("encode", _lib.lib().annf_encode_image_dev(h.h, dev.dptr(img), a.shape[0], a.shape[1], p,
                                                          dev.dptr(buf)))
    got = buf.cpu().numpy()
    assert np.all(got[n * d:] == -7.25)
    pa = O.extract_patches(a, p)
    ref = M.forward(pa, enc.params, p * p * 3, H1, H2, d, bf16_ops=True)
    assert np.abs(got[: n * d].reshape(n, d) - ref).max() <= 3e-3 * np.abs(ref).max()


def test_field_on_encoder_embeddings(gpu_available):
    """Both field entry points with ANNF_PROJ_ENCODER: the plain path equals
    the oracle's rows on the same embeddings, the BF16 path the exact k-NN
    over the natural leaves of the tree built on them."""
    from annfmp_amd import synth, field_on_device, field_mm_on_device
    from annfmp_amd.encoder import PatchEncoder
    a, b = synth.make_pair(synth.small_spec(96, 128), seed=12)
    p, leaf, K = 7, 32, 5
    enc = PatchEncoder(p, DOUT, seed=5)
    pa, pb = O.extract_patches(a, p), O.extract_patches(b, p)
    depth = O.tree_depth(pb.shape[0], leaf)
    nr, nc = a.shape[0] - p + 1, a.shape[1] - p + 1
    res = field_on_device(a, b, p, None, None, depth, K, True, True, projection=enc, return_reduced=True)
    ref = O.OracleTree(res["red_b"], depth)
    final, pos, _ = ref.process_rows(res["red_a"], nr, nc, K, True, True, pa, pb)
    assert np.array_equal(res["nn"], final)
    assert np.array_equal(res["knn"], ref.orig[pos])
    res2 = field_mm_on_device(a, b, p, None, None, depth, K, nprobe=1, precision="bf16", projection=enc,
                              return_reduced=True)
    assert np.array_equal(res2["red_b"].view(np.uint32), res["red_b"].view(np.uint32))
    qleaf = ref.natural_leaves(res2["red_a"])
    wi, wd = O.leaf_knn_exact(pa, pb, qleaf, ref.leaves, ref.orig, 1, K)
    assert np.array_equal(res2["knn"], wi.astype(np.int32))


def test_training_lowers_the_loss(gpu_available):
    """A short run of the HIP training loop on exact-nearest triplets."""
    from annfmp_amd import synth
    from annfmp_amd.encoder import PatchEncoder, make_triplets, train
    h = _handle()
    trip = make_triplets(7, 2, 3072, spec=synth.small_spec(120, 140), seed=77, handle=h)
    assert trip.shape == (6144, 3, 147)
    enc = PatchEncoder(7, DOUT, seed=1)
    train(enc, trip, epochs=3, batch=256, lr=1e-3, handle=h, log_every=8)
    losses = [l for _, l in enc.history]
    print("loss history", [round(l, 4) for l in losses])
    assert np.mean(losses[-3:]) < 0.5 * np.mean(losses[:2])
