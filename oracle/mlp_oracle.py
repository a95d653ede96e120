"""FP64 CPU oracle of the warm-start network (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product never does.  The network has no reference counterpart: it
is the learned warm start of the coefficient fit.

Network (parameter layout of its_mlp_set_params):
    z1 = x W1 + b1,  h1 = gelu(z1)
    z2 = h1 W2 + b2, h2 = gelu(z2)
    z3 = h2 W3 + b3, h3 = gelu(z3)
    y  = h3 . W4 + b4,   loss = mean((y - t)^2)
with gelu(z) = z Phi(z) (erf form).

mimic=True applies bfloat16 rounding at exactly the points where the gfx950
kernels round (MFMA operands: x, W1..W3, h1, h2 and the dZ tiles; W4); every
other operation is binary64.  mimic=False is the plain binary64 function,
which the finite-difference check pins.
"""
import math

import numpy as np

IN, H = 16, 64
NP = IN * H + H + H * H + H + H * H + H + H + 1


def bf16_bits(x):
    """binary32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_round(x):
    return bf16_value(bf16_bits(np.asarray(x, np.float64).astype(np.float32)))


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
                        ("W4", (H,)), ("b4", ())):
        size = int(np.prod(shape)) if shape else 1
        out[name] = p[o:o + size].reshape(shape) if shape else p[o]
        o += size
    assert o == NP
    return out


def pack(d):
    return np.concatenate([np.ravel(d[k]) for k in ("W1", "b1", "W2", "b2", "W3", "b3", "W4")] + [[d["b4"]]])


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p, x, mimic=True):
    P = unpack(p)
    r = bf16_round if mimic else (lambda a: np.asarray(a, np.float64))
    x = r(x)
    W1, W2, W3, W4 = r(P["W1"]), r(P["W2"]), r(P["W3"]), r(P["W4"])
    z1 = x @ W1 + P["b1"]
    h1 = gelu(z1)
    z2 = r(h1) @ W2 + P["b2"]
    h2 = gelu(z2)
    z3 = r(h2) @ W3 + P["b3"]
    h3 = gelu(z3)
    y = h3 @ W4 + P["b4"]
    cache = dict(x=x, W1=W1, W2=W2, W3=W3, W4=W4, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3)
    return y, cache


def loss_and_grad(p, x, t, mimic=True):
    """mean squared error and its gradient (packed like the parameters)."""
    r = bf16_round if mimic else (lambda a: np.asarray(a, np.float64))
    y, c = forward(p, x, mimic)
    n = len(y)
    e = y - np.asarray(t, np.float64)
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = {}
    g["W4"] = c["h3"].T @ dy
    g["b4"] = dy.sum()
    dz3 = np.outer(dy, c["W4"]) * gelu_d(c["z3"])
    g["b3"] = dz3.sum(0)
    g["W3"] = r(c["h2"]).T @ r(dz3)
    dz2 = (r(dz3) @ c["W3"].T) * gelu_d(c["z2"])
    g["b2"] = dz2.sum(0)
    g["W2"] = r(c["h1"]).T @ r(dz2)
    dz1 = (r(dz2) @ c["W2"].T) * gelu_d(c["z1"])
    g["b1"] = dz1.sum(0)
    g["W1"] = c["x"].T @ r(dz1)
    return loss, pack(g), y


def glorot_params(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, (fi, fo) in (("W1", (IN, H)), ("W2", (H, H)), ("W3", (H, H))):
        a = math.sqrt(6.0 / (fi + fo))
        d[name] = rng.uniform(-a, a, size=(fi, fo))
    a = math.sqrt(6.0 / (H + 1))
    d["W4"] = rng.uniform(-a, a, size=H)
    for b in ("b1", "b2", "b3"):
        d[b] = 0.1 * rng.standard_normal(H)
    d["b4"] = 0.05
    return pack(d).astype(np.float32)
