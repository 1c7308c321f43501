"""Pins the FP64 MLP oracle: bf16 rounding spec, GELU, FD gradients, Adam closed form.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from mlp_ref import NP, oracle_grad, unpack

dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))


def test_bf16_round_to_nearest_even(orc):
    L = orc.lib()
    cases = {1.0: 0x3F80, 1.00390625: 0x3F80, 1.01171875: 0x3F82, 1.0078125: 0x3F81, -2.0: 0xC000,
             float("inf"): 0x7F80, 3.3895313892515355e38: 0x7F7F, 3.39e38: 0x7F7F, 3.4e38: 0x7F80, 0.0: 0x0000}
    for v, h in cases.items():
        assert L.orc_f32_to_bf16(v) == h, (v, hex(L.orc_f32_to_bf16(v)))
    assert math.isnan(L.orc_bf16_to_f32(L.orc_f32_to_bf16(float("nan"))))
    assert L.orc_bf16_to_f32(0x3F82) == 1.015625


def test_gelu_closed_form_and_derivative(orc):
    L = orc.lib()
    y, dy = C.c_double(), C.c_double()
    L.orc_gelu(0.0, C.byref(y), C.byref(dy))
    assert y.value == 0.0 and dy.value == pytest.approx(0.5)
    L.orc_gelu(10.0, C.byref(y), None)
    assert y.value == pytest.approx(10.0, rel=1e-12)
    for x in np.linspace(-4, 4, 33):
        h = 1e-6
        yp, ym = C.c_double(), C.c_double()
        L.orc_gelu(x + h, C.byref(yp), None)
        L.orc_gelu(x - h, C.byref(ym), None)
        L.orc_gelu(x, C.byref(y), C.byref(dy))
        assert dy.value == pytest.approx((yp.value - ym.value) / (2 * h), abs=1e-7)


def test_mlp_gradient_finite_difference(orc):
    rng = np.random.default_rng(0)
    prm = rng.standard_normal(NP) * 0.2
    X = rng.standard_normal((5, 22))
    T = rng.standard_normal((5, 3))
    loss, g = oracle_grad(orc, prm, X, T)
    for k in rng.choice(NP, 60, replace=False).tolist() + [5824, 5825, 5826, 0, 1407]:
        h = 1e-6
        pp, pm = prm.copy(), prm.copy()
        pp[k] += h
        pm[k] -= h
        fd = (oracle_grad(orc, pp, X, T)[0] - oracle_grad(orc, pm, X, T)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_adam_closed_form(orc):
    L = orc.lib()
    # f(w) = 0.5 |w - a|^2 ; first bias-corrected step moves every weight by
    # lr * g / (|g| + eps) (m_hat = g, v_hat = g^2)
    a = np.array([1.0, -2.0, 0.5, 3.0])
    w = np.zeros(4)
    m, v = np.zeros(4), np.zeros(4)
    g = w - a
    L.orc_adam_step(dp(w), dp(g), dp(m), dp(v), 4, 0.1, 0.9, 0.999, 1e-8, 1)
    assert np.allclose(w, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    # converges to the minimum
    for step in range(2, 3000):
        g = w - a
        L.orc_adam_step(dp(w), dp(g), dp(m), dp(v), 4, 0.01, 0.9, 0.999, 1e-8, step)
    assert np.allclose(w, a, atol=1e-3)
