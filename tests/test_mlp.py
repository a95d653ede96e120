"""Warm-start network: BF16 MFMA kernels against the FP64 oracle.

The network has no reference counterpart (the reference starts its solve
from zero, ThinShells.cpp:615-617).  The oracle (oracle/mlp_oracle.py) is
pinned by central finite differences of its own binary64 function; the GPU
forward and gradients are compared with the oracle evaluated with bfloat16
rounding at the points where the kernels round (mimic=True), and loosely with
the plain binary64 function.
"""
import numpy as np
import pytest

import mlp_oracle as M


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, M.IN)).astype(np.float32)
    x[:, 15] = 0.0
    xb = M.bf16_bits(x)
    t = (np.sin(M.bf16_value(xb[:, 0])) + 0.3 * M.bf16_value(xb[:, 4]) * M.bf16_value(xb[:, 5])).astype(np.float32)
    return xb, t


def test_oracle_gradient_matches_finite_differences():
    p = M.glorot_params(3).astype(np.float64)
    xb, t = _data(4, 37)
    x = M.bf16_value(xb)
    loss, g, _ = M.loss_and_grad(p, x, t, mimic=False)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(np.arange(0, 1024), 8), rng.choice(np.arange(1024, M.NP), 24), [M.NP - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NP)
        e[i] = h
        lp = M.loss_and_grad(p + e, x, t, mimic=False)[0]
        lm = M.loss_and_grad(p - e, x, t, mimic=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_known_values():
    # gelu at 0 and symmetric points; bf16 rounding of exact and halfway values
    assert M.gelu(np.array([0.0]))[0] == 0.0
    z = np.array([1.3])
    assert abs(M.gelu(z)[0] - M.gelu(-z)[0] - z[0]) < 1e-15     # gelu(z) - gelu(-z) = z
    assert M.bf16_value(M.bf16_bits(np.array([1.0], np.float32)))[0] == 1.0
    # 1 + 2^-8 is halfway between bf16 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16_value(M.bf16_bits(np.array([1 + 2.0 ** -8], np.float32)))[0] == 1.0
    assert M.bf16_value(M.bf16_bits(np.array([1 + 3 * 2.0 ** -8], np.float32)))[0] == 1 + 2.0 ** -6


def test_bf16_rounding_matches_torch():
    import torch
    x = np.random.default_rng(0).standard_normal(4096).astype(np.float32) * 100
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(M.bf16_bits(x), ref)


# ------------------------------------------------------------------ GPU ----
def _dev(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _net(ctx, seed):
    from its_amd import _capi
    net = _capi.Mlp(ctx, seed=0)
    p = M.glorot_params(seed)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    return net, p


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 31, 1000, 20003])
def test_forward_matches_oracle(ctx, n):
    import torch
    print(f"[mlp forward] n={n}", flush=True)
    net, p = _net(ctx, 7)
    xb, _ = _data(8, n)
    dx = _dev(xb.view(np.int16), torch)
    dy = torch.full((n,), float("nan"), device="cuda", dtype=torch.float32)
    net.forward_device(dx.data_ptr(), n, dy.data_ptr())
    torch.cuda.synchronize()
    y = dy.cpu().numpy().astype(np.float64)
    yr, _ = M.forward(p, M.bf16_value(xb), mimic=True)
    scale = np.abs(yr).max()
    assert np.all(np.isfinite(y))
    err = np.abs(y - yr)
    # binary32 vs binary64 pre-activations occasionally round an activation to
    # the neighbouring bfloat16 value (one ulp, 2^-8 relative): rare outliers
    assert np.quantile(err, 0.999) <= 1e-5 * max(scale, 1.0)
    assert err.max() <= 2e-3 * max(scale, 1.0)
    y64, _ = M.forward(p, M.bf16_value(xb), mimic=False)
    assert np.abs(y - y64).max() <= 3e-2 * max(scale, 1.0)
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 777, 9000])
def test_gradient_matches_oracle(ctx, n):
    import torch
    print(f"[mlp grad] n={n}", flush=True)
    net, p = _net(ctx, 9)
    xb, t = _data(10, n)
    dx, dt = _dev(xb.view(np.int16), torch), _dev(t, torch)
    g, loss = net.grad_device(dx.data_ptr(), dt.data_ptr(), n)
    lr, gr, _ = M.loss_and_grad(p, M.bf16_value(xb), t, mimic=True)
    assert abs(loss - lr) <= 1e-4 * lr
    o = 0
    for name, size in (("W1", 1024), ("b1", 64), ("W2", 4096), ("b2", 64), ("W3", 4096), ("b3", 64), ("W4", 64),
                       ("b4", 1)):
        a, b = g[o:o + size].astype(np.float64), gr[o:o + size]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  {name}: rel err {rel:.2e}")
        assert rel <= 2e-3, name
        o += size
    # against the plain binary64 gradient: bfloat16 operands cost about 1%
    _, g64, _ = M.loss_and_grad(p, M.bf16_value(xb), t, mimic=False)
    cos = float(np.dot(g, g64) / (np.linalg.norm(g) * np.linalg.norm(g64)))
    assert cos > 0.999
    net.close()


@pytest.mark.gpu
def test_training_reduces_loss(ctx):
    import torch
    net, _ = _net(ctx, 11)
    xb, t = _data(12, 8192)
    dx, dt = _dev(xb.view(np.int16), torch), _dev(t, torch)
    first = net.train_step(dx.data_ptr(), dt.data_ptr(), len(t), 3e-3)
    for _ in range(300):
        last = net.train_step(dx.data_ptr(), dt.data_ptr(), len(t), 3e-3)
    print("loss", first, "->", last)
    assert last < 0.1 * first
    net.close()
