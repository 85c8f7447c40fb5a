"""CPU restatement of the surrogate MLP's forward and backward passes
(TEST INFRASTRUCTURE ONLY: the checker for neurogpu_amd/csrc/mlp.hip).

Same network and the same BF16 rounding points as the device: inputs,
weights, hidden activations and back-propagated activation gradients are
rounded to bfloat16; every product is accumulated in float64 here (FP32 on the
device), so the two agree to BF16 resolution.  Loss: mean squared error."""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _exact(x):
    return np.asarray(x, dtype=np.float64)


def _f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


def forward(params, x, rounding=True):
    """Network output [n, nout] and the activations a[0..3] (a[0] = input).
    ``rounding=False`` drops every BF16/FP32 rounding (plain FP64), the form
    the tests check against autograd and finite differences."""
    r16, r32 = (bf16, _f32) if rounding else (_exact, _exact)
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(p, dtype=np.float64) for p in params]
    a = [r16(x)]
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        z = a[-1] @ r16(W).T + b
        a.append(r16(np.maximum(z, 0.0)))
    y = a[-1] @ r16(W4).T + b4
    return r32(y), a


def gradients(params, x, t, rounding=True):
    """Gradients of the mean-squared-error loss with respect to
    (W1, b1, W2, b2, W3, b3, W4, b4), and the loss."""
    bf16_, f32_ = (bf16, _f32) if rounding else (_exact, _exact)
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(p, dtype=np.float64) for p in params]
    y, a = forward(params, x, rounding)
    n = y.size
    d = y - t
    loss = float(np.sum(d * d) / n)
    dy = f32_(2.0 * d / n)
    dy16 = bf16_(dy)
    gW4 = dy16.T @ a[3]
    gb4 = dy.sum(axis=0)
    dz = bf16_((dy16 @ bf16_(W4)) * (a[3] > 0))
    grads = [None] * 8
    grads[6], grads[7] = gW4, gb4
    Ws = [W1, W2, W3]
    for l in (2, 1, 0):
        grads[2 * l] = dz.T @ a[l]
        grads[2 * l + 1] = dz.sum(axis=0)
        if l > 0:
            dz = bf16_((dz @ bf16_(Ws[l])) * (a[l] > 0))
    return grads, loss
