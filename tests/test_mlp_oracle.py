"""Pins the FP64 MLP reference (oracle/mlp_oracle.py) by central finite
differences, and its bf16 rounding by known values."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as mo  # noqa: E402


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-3, 65504.0], dtype=np.float32)
    r = mo.bf16_round(x)
    # 1 + 2^-8 is exactly halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == 1.0078125  # rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8
    assert r[5] == 65536.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, mo.NP)
    X = rng.normal(0, 1, (16, mo.F))
    Y = rng.normal(0, 1, 16)
    loss, g = mo.loss_and_grad(w, X, Y)
    idx = rng.choice(mo.NP, 60, replace=False)
    idx = np.concatenate([idx, [mo.NP - 1, mo.H * mo.F, mo.H * mo.F + mo.H]])
    h = 1e-6
    for i in idx:
        wp = w.copy(); wp[i] += h
        wm = w.copy(); wm[i] -= h
        fd = (mo.loss_and_grad(wp, X, Y)[0] - mo.loss_and_grad(wm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_adam_step_closed_forms():
    """Constant gradient: the bias-corrected moments are g and g^2 at every
    step, so each update is lr * g / (|g| + eps); zero gradient: no move; a
    sign flip +1, -1: m_hat_2 = -0.01 / 0.19 and v_hat_2 = 1 exactly."""
    lr = 1e-2
    w0 = np.array([0.5, -0.25, 0.0, 2.0])
    g = np.array([3.0, -0.5, 0.0, 1e-3])
    w, m, v = w0.copy(), np.zeros(4), np.zeros(4)
    for t in range(1, 6):
        w, m, v = mo.adam_step(w, m, v, g, t, lr)
        step = lr * g / (np.abs(g) + mo.EPS)
        assert np.allclose(w, w0 - t * step, rtol=0, atol=1e-12)
    w, m, v = mo.adam_step([1.0], [0.0], [0.0], [1.0], 1, lr)
    w, m, v = mo.adam_step(w, m, v, [-1.0], 2, lr)
    mh, vh = -0.01 / 0.19, 1.0
    assert abs(w[0] - (1.0 - lr / (1 + mo.EPS) - lr * mh / (np.sqrt(vh) + mo.EPS))) < 1e-15
