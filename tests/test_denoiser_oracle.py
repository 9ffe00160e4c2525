"""Pins for the CPU restatement of the denoiser.

* BF16 rounding: hand-worked ties-to-even cases, and agreement with
  torch's float32 -> bfloat16 conversion on random values.
* Backward pass: in float64 (no rounding emulation) the gradient of
  sum(g * pred) must reproduce central finite differences of the forward
  pass, per convolution, so a wrong tap index, transposed weight, missing
  skip connection or pooling factor fails here.
* Layout: the first convolution equals a direct 3x3 sum over taps written
  out per pixel.
"""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import denoiser_ref as dr  # noqa: E402


def test_bf16_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -2.5, 0.0,
                  np.finfo(np.float32).max, np.inf, -np.inf], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0, np.inf, np.inf, -np.inf], np.float32)
    got = dr.bf16(x)
    assert np.array_equal(got[:6], want[:6])
    # the largest float32 lies above the largest finite BF16 + half an ulp: inf
    assert np.isinf(got[6]) and got[6] > 0
    assert np.array_equal(got[7:], want[7:])
    assert np.isnan(dr.bf16(np.array([np.nan], np.float32)))[0]


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(100000) * np.exp(rng.uniform(-30, 30, 100000))).astype(np.float32)
    t = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(dr.bf16(x).view(np.uint32), t.view(np.uint32))


def _random_params(rng, scale=1.0):
    offs, n = dr.param_offsets()
    p = np.zeros(n, np.float64)
    for (cin, cout, _), (wo, bo) in zip(dr.conv_specs(), offs):
        p[wo:wo + cout * 9 * cin] = rng.standard_normal(cout * 9 * cin) * scale * np.sqrt(2.0 / (9 * cin))
        p[bo:bo + cout] = 0.05 * rng.standard_normal(cout)
    return p


def _input(rng, n, h, w):
    x = np.zeros((n, h, w, 9), np.float32)
    x[..., :3] = rng.gamma(1.0, 1.0, (n, h, w, 3))
    x[..., 3:6] = rng.uniform(0, 1, (n, h, w, 3))
    nn = rng.standard_normal((n, h, w, 3))
    x[..., 6:9] = nn / np.linalg.norm(nn, axis=-1, keepdims=True)
    return x


def test_parameter_count():
    _, n = dr.param_offsets()
    # encoder, decoder (concatenated skips) and output layer
    enc = 9 * (16 * 32 + 32 * 32 + 32 * 64 + 64 * 64 + 64 * 96 + 96 * 96 + 96 * 128 + 128 * 128 + 128 * 160
               + 160 * 160) + (32 + 32 + 64 + 64 + 96 + 96 + 128 + 128 + 160 + 160)
    dec = 9 * (288 * 128 + 128 * 128 + 224 * 96 + 96 * 96 + 160 * 64 + 64 * 64 + 96 * 32 + 32 * 32) \
        + (128 + 128 + 96 + 96 + 64 + 64 + 32 + 32)
    out = 9 * 32 * 32 + 32
    assert n == enc + dec + out


def test_first_conv_direct_sum():
    rng = np.random.default_rng(1)
    p = _random_params(rng)
    inp = _input(rng, 1, 16, 16)
    _, cache = dr.forward(p, inp, emulate=False)
    x = dr.pack_input(inp, emulate=False)[0]
    cin, cout, _ = dr.conv_specs()[0]
    w = p[:cout * 9 * cin].reshape(cout, 9, cin)
    b = p[cout * 9 * cin:cout * 9 * cin + cout]
    y = cache["y"][0][0]
    for (yy, xx) in [(0, 0), (7, 3), (15, 15), (0, 15), (9, 0)]:
        acc = b.copy()
        for t in range(9):
            sy, sx = yy + t // 3 - 1, xx + t % 3 - 1
            if 0 <= sy < 16 and 0 <= sx < 16:
                acc += w[:, t, :] @ x[sy, sx]
        assert np.allclose(y[yy, xx], np.maximum(acc, 0.0), rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = _random_params(rng)
    inp = _input(rng, 1, 16, 16)
    pred, cache = dr.forward(p, inp, emulate=False)
    g = rng.standard_normal(pred.shape)
    g[..., 3:] = 0.0
    grads = dr.backward(p, cache, g, emulate=False)
    offs, _ = dr.param_offsets()
    # small enough that no ReLU changes sides within +-eps on this seed
    eps = 1e-7
    for k, ((cin, cout, _), (wo, bo)) in enumerate(zip(dr.conv_specs(), offs)):
        d = np.zeros_like(p)
        d[wo:bo + cout] = rng.standard_normal(bo + cout - wo)
        fp = float((g * dr.forward(p + eps * d, inp, emulate=False)[0]).sum())
        fm = float((g * dr.forward(p - eps * d, inp, emulate=False)[0]).sum())
        fd = (fp - fm) / (2 * eps)
        an = float(grads @ d)
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (k, fd, an)


def test_loss_gradient_is_sign_over_count():
    rng = np.random.default_rng(3)
    pred = rng.standard_normal((1, 4, 4, 32))
    inp = _input(rng, 1, 4, 4)
    tgt = rng.gamma(1.0, 1.0, (1, 4, 4, 3)).astype(np.float32)
    fin = dr.final_output(pred, inp, emulate=False)
    assert np.allclose(fin, pred[..., :3] + np.log1p(inp[..., :3].astype(np.float64)))
    loss, g = dr.loss_grad(fin, tgt, emulate=False)
    d = fin - np.log1p(tgt.astype(np.float64))
    assert np.isclose(loss, np.abs(d).mean())
    assert np.array_equal(g[..., :3] * d.size, np.sign(d)) and not g[..., 3:].any()
