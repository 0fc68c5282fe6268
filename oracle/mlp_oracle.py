"""CPU oracle of the learned patch encoder (FP64 numpy).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the HIP
training and inference kernels; the product package never imports it.

The network (no reference counterpart; it replaces the PCA projection of
annfmp/openmp/base.py:193-200 as the tree's feature map):

    x  = (patch_byte - 128) / 128                       (D inputs)
    Z1 = x W1^T + b1,  H1 = gelu(Z1)                    (h1)
    Z2 = H1 W2^T + b2, H2 = gelu(Z2)                    (h2)
    Z  = H2 W3^T + b3                                   (d outputs)

gelu is the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).  The
loss is PyTorch's TripletMarginLoss(margin, p=2, eps=1e-6, reduction=mean)
over rows [queries; positives; negatives]: mean(max(|q - p + eps| -
|q - n + eps| + margin, 0)).

``bf16=True`` rounds every GEMM operand to BF16 (round to nearest even),
the points where the HIP kernels round; ``bf16=False`` is the plain FP64
network, pinned by finite differences and by torch autograd on the CPU.
"""
from __future__ import annotations

import numpy as np

C0 = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round to the nearest BF16 (ties to even) through FP32, as FP64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = C0 * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C0 * (1.0 + 3 * 0.044715 * z * z)


def split_params(params, D, h1, h2, d):
    params = np.asarray(params, np.float64)
    shapes = [(h1, D), (h1,), (h2, h1), (h2,), (d, h2), (d,)]
    out, o = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(params[o:o + n].reshape(s))
        o += n
    assert o == params.size, "parameter count mismatch"
    return out


def param_count(D, h1, h2, d):
    return h1 * D + h1 + h2 * h1 + h2 + d * h2 + d


def init_params(D, h1, h2, d, seed=0):
    """PyTorch nn.Linear default init: weights and biases U(-1/sqrt(fan_in), +)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((D, h1), (h1, h2), (h2, d)):
        bound = 1.0 / np.sqrt(fan_in)
        parts.append(rng.uniform(-bound, bound, fan_out * fan_in))
        parts.append(rng.uniform(-bound, bound, fan_out))
    return np.concatenate(parts).astype(np.float32)


def inputs(x_u8):
    return (np.asarray(x_u8, np.float64) - 128.0) / 128.0


def forward(x_u8, params, D, h1, h2, d, bf16_ops=True, keep=False):
    W1, b1, W2, b2, W3, b3 = split_params(params, D, h1, h2, d)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    X = inputs(x_u8)
    Z1 = r(X) @ r(W1).T + b1
    H1 = gelu(Z1)
    Z2 = r(H1) @ r(W2).T + b2
    H2 = gelu(Z2)
    Z = r(H2) @ r(W3).T + b3
    if keep:
        return Z, (X, Z1, H1, Z2, H2)
    return Z


def triplet_loss_grad(Z, B, margin, eps=1e-6):
    """Loss and dLoss/dZ for rows [q; p; n] of Z (3B x d)."""
    q, p, n = Z[:B], Z[B:2 * B], Z[2 * B:]
    a = q - p + eps
    b = q - n + eps
    dp = np.sqrt((a * a).sum(1))
    dn = np.sqrt((b * b).sum(1))
    l = dp - dn + margin
    act = (l > 0).astype(np.float64)[:, None]
    loss = np.maximum(l, 0.0).mean()
    ga = a / dp[:, None] * act / B
    gb = b / dn[:, None] * act / B
    dZ = np.concatenate([ga - gb, -ga, gb])
    return loss, dZ


def triplet_grad(x_u8, B, params, D, h1, h2, d, margin, bf16_ops=True):
    """(loss, gradient in the parameter layout) of the triplet loss."""
    W1, b1, W2, b2, W3, b3 = split_params(params, D, h1, h2, d)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    Z, (X, Z1, H1, Z2, H2) = forward(x_u8, params, D, h1, h2, d, bf16_ops, keep=True)
    loss, dZ = triplet_loss_grad(Z, B, margin)
    gW3 = r(dZ).T @ r(H2)
    gb3 = dZ.sum(0)
    dZ2 = (r(dZ) @ r(W3)) * gelu_grad(Z2)
    gW2 = r(dZ2).T @ r(H1)
    gb2 = dZ2.sum(0)
    dZ1 = (r(dZ2) @ r(W2)) * gelu_grad(Z1)
    gW1 = r(dZ1).T @ r(X)
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def adam_step(p, g, m, v, lr, b1, b2, eps, step):
    """One bias-corrected Adam step (FP64); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
