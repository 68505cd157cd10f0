"""Pins the NumPy restatement of the learned-guess network (oracle/unet_oracle.py):
building blocks against textbook routines, the backward pass against finite
differences of the FP64 forward, and Adam's first step against its closed form."""
import numpy as np
import pytest
from scipy import signal

import unet_oracle as U


def test_conv3x3_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(12, 16, 3))
    w = rng.normal(size=(9, 4, 3))
    b = rng.normal(size=4)
    y = U.conv3x3(x, w, b)
    for co in range(4):
        ref = b[co] + sum(signal.correlate2d(x[:, :, ci], w[:, co, ci].reshape(3, 3), mode="same")
                          for ci in range(3))
        assert np.allclose(y[:, :, co], ref, atol=1e-12)


def test_conv_gradients_are_adjoints():
    """<conv(x), dz> = <x, conv_dx(dz)> and = <w, conv_dw(x, dz)> (bias-free)."""
    rng = np.random.default_rng(1)
    x = rng.normal(size=(8, 12, 5))
    w = rng.normal(size=(9, 6, 5))
    dz = rng.normal(size=(8, 12, 6))
    y = U.conv3x3(x, w, np.zeros(6))
    lhs = (y * dz).sum()
    assert lhs == pytest.approx((x * U.conv3x3_dx(dz, w)).sum(), rel=1e-12)
    assert lhs == pytest.approx((w * U.conv3x3_dw(x, dz)).sum(), rel=1e-12)


def test_pool_and_upsample():
    x = np.arange(4 * 6 * 2, dtype=float).reshape(4, 6, 2)
    p = U.pool2(x)
    assert np.allclose(p, x.reshape(2, 2, 3, 2, 2).mean(axis=(1, 3)))
    u = U.up2(p)
    assert u.shape == x.shape and np.allclose(u[::2, ::2], p) and np.allclose(u[1::2, 1::2], p)


def test_bf16_rounding_ties_to_even():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 1e-3], np.float32)
    r = U.bf16(v)
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6]
    assert r[4] == -2.5 and abs(r[5] - 1e-3) <= 1e-3 * 2**-8


def test_layout_parameter_count():
    _, hw, hb, n = U.layout()
    convs = sum(9 * co * ci + co for (_, _, ci, _, _, _, co) in U.CONVS)
    assert hw == convs and hb == convs + 16 and n == convs + 17


@pytest.mark.slow
def test_backward_matches_finite_differences():
    """FP64 (no BF16 rounding) gradient against central differences, on
    parameters of every layer."""
    rng = np.random.default_rng(3)
    N = 16
    _, _, _, n = U.layout()
    params = rng.normal(0, 0.15, n)
    blocks, hw, hb, _ = U.layout()
    params[hw:hw + 16] = rng.normal(0, 0.5, 16)
    x = rng.normal(size=(N, N, 2))
    tgt = rng.normal(size=(N, N))
    mask = (rng.random((N, N)) > 0.2).astype(float)
    loss, g = U.loss_and_grads(params, x, tgt, mask, bf=False)
    picks = [hw + 3, hb]
    for (w, b, ci, co) in blocks:
        picks += [w + int(rng.integers(0, 9 * co * min(ci, 2))), b + int(rng.integers(0, co))]
    eps = 1e-6
    bad = []
    for k in picks:
        p1 = params.copy(); p1[k] += eps
        p2 = params.copy(); p2[k] -= eps
        fd = (U.loss_and_grads(p1, x, tgt, mask, bf=False)[0] - U.loss_and_grads(p2, x, tgt, mask, bf=False)[0]) / (2 * eps)
        if abs(fd - g[k]) > 1e-5 * max(1.0, abs(fd)) + 1e-7:
            bad.append((k, fd, g[k]))
    assert not bad, bad


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -4.0, 1e-3], np.float32)
    m = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    U.adam(p, g, m, v, 1, 1e-2)
    assert np.allclose(p, [1.0 - 1e-2, -2.0 + 1e-2, 0.5 - 1e-2], atol=1e-6)


@pytest.mark.slow
def test_residual_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    N = 16
    _, hw, hb, n = U.layout()
    params = rng.normal(0, 0.15, n)
    params[hw:hw + 16] = rng.normal(0, 0.5, 16)
    x = rng.normal(size=(N, N, 2))
    r = rng.normal(size=(N, N))
    mask = np.ones((N, N))
    mask[0, :] = mask[:, 0] = mask[-1, :] = 0
    cxd = cyd = 0.25
    loss, g = U.residual_loss_and_grads(params, x, r, mask, cxd, cyd, bf=False)
    bad = []
    for k in [hw + 1, hb, 5, 200, n - 100]:
        p1 = params.copy(); p1[k] += 1e-6
        p2 = params.copy(); p2[k] -= 1e-6
        fd = (U.residual_loss_and_grads(p1, x, r, mask, cxd, cyd, bf=False)[0]
              - U.residual_loss_and_grads(p2, x, r, mask, cxd, cyd, bf=False)[0]) / 2e-6
        if abs(fd - g[k]) > 1e-5 * max(1.0, abs(fd)) + 1e-8:
            bad.append((k, fd, g[k]))
    assert not bad, bad
