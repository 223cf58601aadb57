"""Pins of the surrogate-network restatement (oracle/mlp_oracle.py) on the CPU.

  bf16()                 against torch's float32 -> bfloat16 conversion (round to nearest even)
  forward()              a one-path chain network against its closed form (pins the flat layout,
                         the input normalisation and which rows/columns are inputs and outputs)
  loss_and_gradient()    against torch autograd on the same float64 network and against central
                         finite differences
  BF16 emulation         close to the float64 network (the roundings are small perturbations)
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as mo

torch = pytest.importorskip("torch")


def _inputs(n, seed=0):
    rng = np.random.default_rng(seed)
    inp = np.empty((n, 5), np.float32)
    inp[:, :3] = rng.normal(0.0, 0.01, size=(n, 3))
    inp[:, 3] = rng.uniform(25.0, 40.0, size=n)
    inp[:, 4] = rng.uniform(1.5, 3.5, size=n)
    return inp


def _weights(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, scale, size=mo.NPARAM).astype(np.float32)
    w[~mo.active_mask()] = 0.0
    return w


def test_layout_sizes():
    assert mo.NPARAM == 55824
    assert mo.active_mask().sum() == (128 * 5 + 128) + 3 * (128 * 128 + 128) + (3 * 128 + 3)


def test_bf16_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-30, 1000), [0.0, -0.0, 1.0 + 2 ** -8,
                        1.0 + 3 * 2 ** -8, 65504.0, 3.4e38, np.inf, -np.inf]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = mo.bf16(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert got[-6] == 1.0 and got[-5] == np.float32(1.0 + 2 ** -6)  # ties go to even


def test_chain_network_closed_form():
    # only one path through the network: input phi -> hidden unit 7 -> 11 -> 2 -> 100 -> output 1
    w = np.zeros(mo.NPARAM, np.float32)
    path = [(7, 3), (11, 7), (2, 11), (100, 2), (1, 100)]
    c = [0.9, -1.3, 0.7, 1.1, 2.0]
    bias = [0.1, -0.2, 0.05, 0.3, -0.4]
    for l, ((o, k), cl, bl) in enumerate(zip(path, c, bias)):
        w[mo.weight_offset(l) + o * mo.K[l] + k] = cl
        w[mo.bias_offset(l) + o] = bl
    inp = np.array([[0.01, -0.02, 0.005, 36.5, 2.9]], np.float64)
    x3 = (36.5 - 32.5) / 7.5
    silu = lambda z: z / (1.0 + math.exp(-z))
    a = silu(0.9 * x3 + 0.1)
    a = silu(-1.3 * a - 0.2)
    a = silu(0.7 * a + 0.05)
    a = silu(1.1 * a + 0.3)
    y1 = 2.0 * a - 0.4
    # the other outputs see only the zero biases of layer 5 (b_5[0] and b_5[2] are 0)
    y, _ = mo.forward(w, inp, emulate_bf16=False)
    assert y.shape == (1, 3)
    assert abs(y[0, 1] - y1) < 1e-6 and y[0, 0] == 0.0 and y[0, 2] == 0.0
    H = mo.predict(w, inp, emulate_bf16=False)
    np.testing.assert_allclose(H[0], inp[0, :3] + 1e-3 * np.array([0.0, y1, 0.0]), rtol=0, atol=1e-9)


def _torch_grad(w, inp, H):
    layers = mo.unpack(w)
    x = torch.from_numpy(mo.normalise(inp).astype(np.float64))
    params = []
    for W, b in layers:
        params.append((torch.tensor(W, requires_grad=True), torch.tensor(b, requires_grad=True)))
    a = x
    for l in range(4):
        a = torch.nn.functional.silu(a @ params[l][0].T + params[l][1])
    y = (a @ params[4][0].T + params[4][1])[:, :3]
    t = torch.from_numpy(mo.targets(inp, H))
    loss = ((y - t) ** 2).sum() / y.shape[0]
    loss.backward()
    g = np.zeros(mo.NPARAM)
    for l, (W, b) in enumerate(params):
        g[mo.weight_offset(l):mo.bias_offset(l)] = W.grad.numpy().ravel()
        g[mo.bias_offset(l):mo.bias_offset(l) + mo.N[l]] = b.grad.numpy()
    g[~mo.active_mask()] = 0.0
    return float(loss), g


def test_gradient_matches_torch_autograd():
    w = _weights(2, 0.15)
    inp = _inputs(37, 3)
    H = inp[:, :3] + 1e-3 * np.random.default_rng(4).normal(0, 1, size=(37, 3))
    l0, g0 = mo.loss_and_gradient(w, inp, H, emulate_bf16=False)
    l1, g1 = _torch_grad(w, inp, H)
    assert abs(l0 - l1) <= 1e-12 * max(1.0, l1)
    assert np.max(np.abs(g0 - g1)) <= 1e-10 * max(1.0, np.max(np.abs(g1)))


def test_gradient_finite_differences():
    w = _weights(5, 0.15).astype(np.float64)
    inp = _inputs(9, 6)
    H = inp[:, :3] + 1e-3 * np.random.default_rng(7).normal(0, 1, size=(9, 3))
    _, g = mo.loss_and_gradient(w, inp, H, emulate_bf16=False)
    rng = np.random.default_rng(8)
    act = np.flatnonzero(mo.active_mask())
    picks = []
    for l in range(5):  # a few weights and the biases of every layer
        lo, hi = mo.weight_offset(l), mo.bias_offset(l) + mo.N[l]
        cand = act[(act >= lo) & (act < hi)]
        picks += list(rng.choice(cand, 6, replace=False))
    for i in picks:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        lp, _ = mo.loss_and_gradient(wp, inp, H, emulate_bf16=False)
        lm, _ = mo.loss_and_gradient(wm, inp, H, emulate_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_emulation_is_a_small_perturbation():
    w = _weights(9, 0.1)
    inp = _inputs(200, 10)
    y0, _ = mo.forward(w, inp, emulate_bf16=False)
    y1, _ = mo.forward(w, inp, emulate_bf16=True)
    assert np.max(np.abs(y1 - y0)) <= 0.03 * np.max(np.abs(y0))
    H = inp[:, :3] + 1e-3 * y0 * 0.5
    l0, g0 = mo.loss_and_gradient(w, inp, H, emulate_bf16=False)
    l1, g1 = mo.loss_and_gradient(w, inp, H, emulate_bf16=True)
    cos = g0 @ g1 / (np.linalg.norm(g0) * np.linalg.norm(g1))
    assert cos > 0.99


def test_zero_residual_gives_zero_gradient():
    w = _weights(11, 0.1)
    inp = _inputs(16, 12)
    H = mo.predict(w, inp, emulate_bf16=False)
    loss, g = mo.loss_and_gradient(w, inp, H, emulate_bf16=False)
    assert loss < 1e-20 and np.max(np.abs(g)) < 1e-9
