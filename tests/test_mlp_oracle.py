"""Pins of the surrogate's CPU restatement (oracle/mlp_ref.py), on the CPU.

With its BF16/FP32 roundings switched off, mlp_ref.gradients must be the
exact gradient of the mean-squared-error loss of the 12 -> 256 -> 256 -> 256
-> 20 ReLU network: it is checked against torch autograd in float64 and
against central finite differences of the FP64 loss.  Plausible mistakes
planted in a copy of the oracle's source (a ReLU mask taken from the wrong
layer) must fail the same check.  With the roundings on, the oracle is the
same function to BF16 resolution."""
import inspect

import numpy as np
import pytest

from oracle import mlp_ref

NIN, H, NOUT, NS = 12, 256, 20, 48


def _net(seed=0):
    rng = np.random.default_rng(seed)
    shapes = [(H, NIN), (H,), (H, H), (H,), (H, H), (H,), (NOUT, H), (NOUT,)]
    params = []
    for s in shapes:
        fan = s[1] if len(s) == 2 else H
        params.append(rng.normal(0, np.sqrt(2.0 / fan), size=s) if len(s) == 2 else rng.normal(0, 0.1, size=s))
    x = rng.uniform(-1, 1, size=(NS, NIN))
    t = rng.normal(size=(NS, NOUT))
    return params, x, t


def _torch_grads(params, x, t):
    torch = pytest.importorskip("torch")
    P = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in params]
    a = torch.tensor(x, dtype=torch.float64)
    for W, b in ((P[0], P[1]), (P[2], P[3]), (P[4], P[5])):
        a = torch.relu(a @ W.T + b)
    y = a @ P[6].T + P[7]
    loss = ((y - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    return [p.grad.numpy() for p in P], float(loss)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (np.linalg.norm(b) + 1e-300))


def _worst_against_autograd(grad_fn, seed=0):
    params, x, t = _net(seed)
    g, loss = grad_fn(params, x, t, rounding=False)
    ref, rloss = _torch_grads(params, x, t)
    assert abs(loss - rloss) <= 1e-12 * abs(rloss)
    return max(_rel(a, b) for a, b in zip(g, ref))


def test_exact_gradients_equal_autograd():
    for seed in (0, 1, 2):
        assert _worst_against_autograd(mlp_ref.gradients, seed) < 1e-12


def test_exact_gradients_equal_finite_differences():
    params, x, t = _net(3)
    g, _ = mlp_ref.gradients(params, x, t, rounding=False)
    rng = np.random.default_rng(4)
    h = 1e-6

    def loss(ps):
        y, _ = mlp_ref.forward(ps, x, rounding=False)
        return float(np.mean((y - t) ** 2))

    for k, p in enumerate(params):
        for _ in range(6):
            idx = tuple(int(rng.integers(n)) for n in p.shape)
            up = [q.copy() for q in params]
            dn = [q.copy() for q in params]
            up[k][idx] += h
            dn[k][idx] -= h
            fd = (loss(up) - loss(dn)) / (2 * h)
            assert abs(fd - g[k][idx]) <= 1e-7 + 1e-5 * abs(fd), (k, idx, fd, g[k][idx])


@pytest.mark.parametrize("wrong", [("(a[l] > 0)", "(a[l + 1] > 0)"), ("(a[3] > 0)", "(a[2] > 0)"),
                                   ("gb4 = dy.sum(axis=0)", "gb4 = dy.sum(axis=0) * 0.5")])
def test_planted_mistakes_fail(wrong):
    src = inspect.getsource(mlp_ref.gradients)
    assert wrong[0] in src
    ns = dict(vars(mlp_ref))
    exec(src.replace(wrong[0], wrong[1]), ns)
    assert _worst_against_autograd(ns["gradients"]) > 1e-3


def test_rounded_oracle_is_the_same_function():
    """The BF16-rounded gradients (what the device kernels compute) differ
    from the exact ones only by rounding: each BF16 rounding is 2^-9
    relative, and back-propagation through three rounded layers (plus ReLU
    masks flipped where a rounded activation crosses zero) leaves about 5 %
    in the first layer's weight gradient."""
    params, x, t = _net(5)
    g16, l16 = mlp_ref.gradients(params, x, t)
    g64, l64 = mlp_ref.gradients(params, x, t, rounding=False)
    assert abs(l16 - l64) < 2e-2 * l64
    for a, b in zip(g16, g64):
        assert _rel(a, b) < 1e-1
