"""The numpy statement of the warm-start network (oracle/unet_oracle.py),
pinned on the CPU: bf16 rounding on known values, the patch convolution
against a direct loop, the adjoint identities of the input and weight
gradients, the whole network's backward pass against central finite
differences in float64, and one Adam step against its closed form."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import unet_oracle as U  # noqa: E402


class F64:
    round = staticmethod(lambda a: a)


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5,
                     np.inf, 1e-40], dtype=np.float32)
    out = U.bf16(vals)
    # ties to even: 1 + ulp/2 -> 1, 1 + 3 ulp/2 -> 1 + 2 ulp; below half -> down
    assert out[0] == 1.0 and out[1] == 1.0 and out[2] == 1.0 + 2 * ulp and out[3] == 1.0
    assert out[4] == -2.5
    assert np.isinf(out[5])
    assert np.isnan(U.bf16(np.array([np.nan], dtype=np.float32)))[0]
    # every output has a zero low half
    assert (out.view(np.uint32) & 0xFFFF == 0).all()


def direct_conv(x, w, b):
    """Loops: zero rows beyond the grid, periodic columns."""
    n, h, wd, cin = x.shape
    cout = w.shape[0]
    out = np.zeros((n, h, wd, cout))
    for s in range(n):
        for y in range(h):
            for xx in range(wd):
                acc = b.astype(np.float64).copy()
                for ky in range(3):
                    yy = y + ky - 1
                    if yy < 0 or yy >= h:
                        continue
                    for kx in range(3):
                        col = (xx + kx - 1) % wd
                        acc += w[:, ky, kx, :] @ x[s, yy, col, :]
                out[s, y, xx] = acc
    return out


def test_patch_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 6, 3))
    w = rng.standard_normal((5, 3, 3, 3))
    b = rng.standard_normal(5)
    ref = direct_conv(x, w, b)
    p = U.patches(x)
    got = (p.reshape(-1, 27) @ w.reshape(5, -1).T + b).reshape(ref.shape)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    # forward applies ReLU after the same sum
    np.testing.assert_allclose(U.conv_forward(x, w, b, F64), np.maximum(ref, 0.0),
                               rtol=1e-12, atol=1e-12)


def test_gradient_adjoint_identities():
    """<conv(x, w), g> = <x, input_grad(g, w)> = <w, weight_grad(x, g)>."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 8, 4))
    w = rng.standard_normal((3, 3, 3, 4))
    g = rng.standard_normal((2, 6, 8, 3))
    p = U.patches(x)
    y = (p.reshape(-1, 36) @ w.reshape(3, -1).T).reshape(g.shape)
    lhs = float((y * g).sum())
    assert abs(float((x * U.conv_input_grad(g, w)).sum()) - lhs) <= 1e-10 * abs(lhs)
    assert abs(float((w * U.conv_weight_grad(x, g)).sum()) - lhs) <= 1e-10 * abs(lhs)


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    params = {k: v.astype(np.float64) for k, v in U.init_params(3).items()}
    x = rng.standard_normal((2, 8, 16, U.IN_CHANNELS))
    target = rng.standard_normal((2, 8, 16))

    def loss_of(p):
        out, _ = U.forward(p, x, fmt=F64, dtype=np.float64)
        return U.loss_and_grad_out(out, target)[0]

    out, acts = U.forward(params, x, fmt=F64, dtype=np.float64)
    _, g_out = U.loss_and_grad_out(out, target)
    grads = U.backward(params, acts, g_out, fmt=F64, dtype=np.float64)
    assert set(grads) == set(params)
    # eps small enough that no ReLU input crosses zero along the direction
    # (at 1e-5 some do and the difference quotient is off by ~1e-3)
    eps = 1e-7
    for trial in range(3):
        d = {k: rng.standard_normal(v.shape) for k, v in params.items()}
        plus = {k: params[k] + eps * d[k] for k in params}
        minus = {k: params[k] - eps * d[k] for k in params}
        fd = (loss_of(plus) - loss_of(minus)) / (2 * eps)
        an = sum(float((grads[k] * d[k]).sum()) for k in params)
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-8, (trial, fd, an)
    # one direction through each layer alone, so no layer can hide behind
    # another's contribution
    for name in ("e0a.w", "e3.w", "d2a.w", "d0.b", "head.w"):
        d = rng.standard_normal(params[name].shape)
        p1 = dict(params)
        p2 = dict(params)
        p1[name] = params[name] + eps * d
        p2[name] = params[name] - eps * d
        fd = (loss_of(p1) - loss_of(p2)) / (2 * eps)
        an = float((grads[name] * d).sum())
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-8, (name, fd, an)


def test_adam_first_step_closed_form():
    p = {"a": np.array([1.0, -2.0, 0.5], dtype=np.float32)}
    g = {"a": np.array([0.3, -4.0, 1e-3], dtype=np.float32)}
    state = {}
    U.adam_step(p, g, state, lr=0.01)
    # first step: m/c1 = g, v/c2 = g^2, so the step is lr g / (|g| + eps)
    expect = np.array([1.0, -2.0, 0.5]) - 0.01 * g["a"] / (np.abs(g["a"]) + 1e-8)
    np.testing.assert_allclose(p["a"], expect, rtol=1e-6)
