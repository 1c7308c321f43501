"""FP64 references of the warm-start MLP used by the tests (not product code)."""
import ctypes as C

import numpy as np

NP = 5827
SL = {"W1": (0, (64, 22)), "b1": (1408, (64,)), "W2": (1472, (64, 64)), "b2": (5568, (64,)),
      "W3": (5632, (3, 64)), "b3": (5824, (3,))}


def unpack(prm):
    return {k: np.asarray(prm[o:o + int(np.prod(s))], np.float64).reshape(s) for k, (o, s) in SL.items()}


def bf16(a):
    """Round-to-nearest-even to bf16, returned as float64 values."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    c = 0.7978845608028654
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def forward_bf16(prm, X):
    """The kernel's rounding points: bf16 weights, inputs and activations; wide accumulation."""
    P = unpack(prm)
    W1, W2, W3 = bf16(P["W1"].astype(np.float32)), bf16(P["W2"].astype(np.float32)), bf16(P["W3"].astype(np.float32))
    x = bf16(X)
    a1 = bf16(gelu(x @ W1.T + P["b1"]).astype(np.float32))
    a2 = bf16(gelu(a1 @ W2.T + P["b2"]).astype(np.float32))
    return a2 @ W3.T + P["b3"]


def oracle_grad(orc, prm, X, T):
    """FP64 oracle loss and gradient (weights as given, no bf16 rounding)."""
    L = orc.lib()
    P = unpack(np.asarray(prm, np.float64))
    args = [np.ascontiguousarray(P[k], np.float64) for k in ("W1", "b1", "W2", "b2", "W3", "b3")]
    g = [np.zeros_like(a) for a in args]
    x = np.ascontiguousarray(X, np.float64)
    t = np.ascontiguousarray(T, np.float64)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    loss = L.orc_mlp_loss_grad(*[dp(a) for a in args], dp(x), dp(t), len(x), *[dp(a) for a in g])
    flat = np.zeros(NP)
    for k, a in zip(("W1", "b1", "W2", "b2", "W3", "b3"), g):
        o, s = SL[k]
        flat[o:o + a.size] = a.reshape(-1)
    return loss, flat
