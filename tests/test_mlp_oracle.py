"""CPU checks of the encoder oracle (oracle/mlp_oracle.py): its backward
against central finite differences, its loss and gradients against torch
autograd on the CPU (TripletMarginLoss, tanh GELU), its BF16 rounding, and
its Adam step against torch.optim.Adam."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _tiny(seed=0, D=12, h1=7, h2=5, d=3, B=6):
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.6, M.param_count(D, h1, h2, d))
    x = rng.integers(0, 256, (3 * B, D)).astype(np.uint8)
    return params, x, (D, h1, h2, d), B


def test_backward_matches_finite_differences():
    params, x, dims, B = _tiny()
    loss, g = M.triplet_grad(x, B, params, *dims, margin=0.4, bf16_ops=False)
    assert loss > 0
    num = np.empty_like(params)
    h = 1e-6
    for i in range(params.size):
        e = np.zeros_like(params)
        e[i] = h
        lp, _ = M.triplet_grad(x, B, params + e, *dims, margin=0.4, bf16_ops=False)
        lm, _ = M.triplet_grad(x, B, params - e, *dims, margin=0.4, bf16_ops=False)
        num[i] = (lp - lm) / (2 * h)
    assert np.abs(num - g).max() <= 1e-6 * max(1.0, np.abs(g).max())


def test_loss_and_gradients_match_torch_autograd():
    torch = pytest.importorskip("torch")
    params, x, (D, h1, h2, d), B = _tiny(seed=3, B=9)
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a, dtype=torch.float64, requires_grad=True)
                              for a in M.split_params(params, D, h1, h2, d)]
    X = torch.tensor(M.inputs(x))
    F = torch.nn.functional
    H1 = F.gelu(F.linear(X, W1, b1), approximate="tanh")
    H2 = F.gelu(F.linear(H1, W2, b2), approximate="tanh")
    Z = F.linear(H2, W3, b3)
    loss = F.triplet_margin_loss(Z[:B], Z[B:2 * B], Z[2 * B:], margin=0.3, p=2, eps=1e-6)
    loss.backward()
    want = np.concatenate([t.grad.numpy().ravel() for t in (W1, b1, W2, b2, W3, b3)])
    lo, g = M.triplet_grad(x, B, params, D, h1, h2, d, margin=0.3, bf16_ops=False)
    assert abs(lo - loss.item()) <= 1e-12
    assert np.abs(g - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_bf16_rounding():
    # ties go to the even mantissa; above a tie rounds up
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 255.0, 257.0, 259.0])
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 255.0, 256.0, 260.0])
    assert np.array_equal(M.bf16(vals), want)
    assert np.array_equal(M.bf16(np.arange(-128, 128) / 128.0), np.arange(-128, 128) / 128.0)


def test_adam_step_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(7)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    t = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for s, g in enumerate(grads, 1):
        t.grad = torch.tensor(g)
        opt.step()
        p, m, v = M.adam_step(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, s)
    assert np.abs(p - t.detach().numpy()).max() <= 1e-12


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    num = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(num - M.gelu_grad(z)).max() <= 1e-8
