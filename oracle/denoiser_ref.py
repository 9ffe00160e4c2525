"""CPU restatement of the learned denoiser (test infrastructure only).

Only tests/, the package smoke check and bench.py's CPU-baseline leg use
this module, and only as the oracle of the HIP kernels in
sundog_amd/csrc/denoise.hip; nothing shipped imports it.

The reference denoises through the vendor AI denoiser
(src/denoise.cpp:18-53: HDR beauty + albedo + normal guides in, denoised
HDR out), whose network and arithmetic are closed. The interface is kept
(same guides, same HDR in/out); the network is this project's own U-Net,
so parity here is against this restatement, which is itself pinned by
finite differences (tests/test_denoiser_oracle.py): the backward pass must
reproduce directional derivatives of the forward pass in float64.

Network, parameter layout and rounding points follow the kernels:
  * input: bf16(log1p(max(beauty, 0))), bf16(albedo), bf16(normal), zero
    padding to 16 channels;
  * 3x3 convolutions, weights [cout][tap][cin] with tap = 3*(dy+1)+(dx+1),
    then bias[cout]; BF16 weights, FP32 (here float64) accumulation, bias
    added after accumulation, ReLU, BF16 store of hidden activations; the
    output convolution (cout padded to 32, 3 used) stores FP32;
  * 2x2 average pooling of BF16 values, rounded to BF16; nearest 2x
    upsampling concatenated with the encoder skip (upsampled part first);
  * final output (log space): the network's first 3 output channels plus
    log1p(max(beauty, 0)) in float32 (a residual on the noisy input);
  * loss: mean |final - log1p(max(target, 0))| over the 3 channels;
  * backward: upstream gradients are rounded to BF16 where the kernels feed
    them to MFMA (backward-data and backward-weights), bias gradients and
    data gradients stay unrounded.
With emulate=False every rounding is skipped and the arithmetic is float64.
"""

import numpy as np

CH = (32, 64, 96, 128, 160)
IN_PAD = 16
OUT_PAD = 32


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as
    float32. Inf/NaN pass through by truncation, as in the kernels."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).reshape(a.shape)


def conv_specs():
    """(cin, cout, relu) per convolution in execution order."""
    specs = []
    prev = IN_PAD
    for c in CH:
        specs += [(prev, c, True), (c, c, True)]
        prev = c
    for l in range(3, -1, -1):
        specs += [(CH[l + 1] + CH[l], CH[l], True), (CH[l], CH[l], True)]
    specs.append((CH[0], OUT_PAD, False))
    return specs


def param_offsets():
    offs, n = [], 0
    for cin, cout, _ in conv_specs():
        w = n
        n += cout * 9 * cin
        offs.append((w, n))
        n += cout
    return offs, n


def _w(params, k):
    cin, cout, _ = conv_specs()[k]
    wo, bo = param_offsets()[0][k]
    return params[wo:wo + cout * 9 * cin].reshape(cout, 9 * cin), params[bo:bo + cout]


def _cols(x):
    """[N,H,W,C] -> [N*H*W, 9*C] with zero padding, tap-major then channel."""
    n, h, w, c = x.shape
    xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
    xp[:, 1:-1, 1:-1] = x
    taps = [xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    return np.stack(taps, axis=3).reshape(n * h * w, 9 * c)


def _uncols(dcols, shape):
    """Adjoint of _cols: scatter-add tap gradients back to [N,H,W,C]."""
    n, h, w, c = shape
    d = dcols.reshape(n, h, w, 9, c)
    xp = np.zeros((n, h + 2, w + 2, c), dcols.dtype)
    t = 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            xp[:, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w] += d[:, :, :, t]
            t += 1
    return xp[:, 1:-1, 1:-1]


def pack_input(inp, emulate=True):
    """[N,H,W,9] -> [N,H,W,16] network input."""
    inp = np.asarray(inp, np.float32)
    n, h, w, _ = inp.shape
    x = np.zeros((n, h, w, IN_PAD), np.float64 if not emulate else np.float32)
    x[..., :3] = np.log1p(np.maximum(inp[..., :3], 0.0).astype(np.float32 if emulate else np.float64))
    x[..., 3:9] = inp[..., 3:9]
    return bf16(x) if emulate else x.astype(np.float64)


def _round(x, emulate):
    return bf16(x.astype(np.float32)).astype(np.float64) if emulate else x


def _pool(x, emulate):
    n, h, w, c = x.shape
    x32 = x.astype(np.float32) if emulate else x
    s = ((x32[:, 0::2, 0::2] + x32[:, 0::2, 1::2]) + x32[:, 1::2, 0::2]) + x32[:, 1::2, 1::2]
    out = (np.float32(0.25) if emulate else 0.25) * s
    return _round(out.astype(np.float64), emulate)


def _up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def forward(params, inp, emulate=True):
    """Returns (pred [N,H,W,32] float64, cache)."""
    params = np.asarray(params, np.float32 if emulate else np.float64)
    specs = conv_specs()
    x0 = pack_input(inp, emulate).astype(np.float64)
    cache = {"x_in": {}, "y": {}, "pooled": {}, "cat": {}, "shape": x0.shape}

    def conv(k, x):
        cin, cout, relu = specs[k]
        w, b = _w(params, k)
        wq = bf16(w).astype(np.float64) if emulate else w.astype(np.float64)
        cols = _cols(x)
        acc = cols @ wq.T
        if emulate:
            acc = acc.astype(np.float32).astype(np.float64)
        acc = acc + b.astype(np.float64)
        if relu:
            acc = np.maximum(acc, 0.0)
        n, h, w_, _ = x.shape
        out = acc.reshape(n, h, w_, cout)
        cache["x_in"][k] = x
        if k < len(specs) - 1:
            out = _round(out, emulate)
        else:
            out = out.astype(np.float32).astype(np.float64) if emulate else out
        cache["y"][k] = out
        return out

    cur = x0
    k = 0
    for l in range(5):
        if l > 0:
            cur = _pool(cache["y"][k - 1], emulate)
            cache["pooled"][l - 1] = cur
        cur = conv(k, cur)
        cur = conv(k + 1, cur)
        k += 2
    for l in range(3, -1, -1):
        cat = np.concatenate([_up(cur), cache["y"][2 * l + 1]], axis=3)
        cache["cat"][l] = cat
        cur = conv(k, cat)
        cur = conv(k + 1, cur)
        k += 2
    pred = conv(k, cur)
    return pred, cache


def final_output(pred, inp, emulate=True):
    """Network output [N,H,W,32] -> final log-space image [N,H,W,3]."""
    b = np.maximum(np.asarray(inp, np.float32)[..., :3], 0.0)
    if emulate:
        return (pred[..., :3].astype(np.float32) + np.log1p(b)).astype(np.float64)
    return pred[..., :3] + np.log1p(b.astype(np.float64))


def loss_grad(final, target, emulate=True):
    """L1 in log space: (loss, dL/dnet [N,H,W,32]) for the final output
    [N,H,W,3] (d final / d net = 1 on the 3 used channels)."""
    t = np.log1p(np.maximum(np.asarray(target, np.float32), 0.0)).astype(np.float64)
    if emulate:
        d = (final.astype(np.float32) - t.astype(np.float32)).astype(np.float64)
    else:
        d = final - t
    cnt = d.size
    g = np.zeros(final.shape[:3] + (OUT_PAD,))
    g[..., :3] = np.sign(d) / cnt
    if emulate:
        g = g.astype(np.float32).astype(np.float64)
    return float(np.abs(d).sum() / cnt), g


def backward(params, cache, gpred, emulate=True):
    """Gradient of sum(gpred * pred) w.r.t. the flat parameter vector."""
    params = np.asarray(params, np.float32 if emulate else np.float64)
    specs = conv_specs()
    offs, n_params = param_offsets()
    grads = np.zeros(n_params, np.float64)
    gy = {}

    def conv_back(k, dy, want_dx=True):
        cin, cout, relu = specs[k]
        y = cache["y"][k]
        if relu:
            dy = np.where(y > 0.0, dy, 0.0)
        x = cache["x_in"][k]
        n, h, w_, _ = x.shape
        dyf = dy.reshape(-1, cout)
        dyq = _round(dyf, emulate)
        cols = _cols(x)
        wo, bo = offs[k]
        grads[wo:wo + cout * 9 * cin] = (dyq.T @ cols).reshape(-1)
        grads[bo:bo + cout] = dyf.sum(axis=0)
        if not want_dx:
            return None
        w, _ = _w(params, k)
        wq = bf16(w).astype(np.float64) if emulate else w.astype(np.float64)
        dx = _uncols(dyq @ wq, (n, h, w_, cin))
        if emulate:
            dx = dx.astype(np.float32).astype(np.float64)
        return dx

    nc = len(specs)
    k = nc - 1
    g = conv_back(k, gpred)
    k -= 1
    enc_extra = {}
    for l in range(4):
        g = conv_back(k, g)
        k -= 1
        gcat = conv_back(k, g)
        k -= 1
        ca = CH[l + 1]
        gu = gcat[..., :ca]
        g = ((gu[:, 0::2, 0::2] + gu[:, 0::2, 1::2]) + gu[:, 1::2, 0::2]) + gu[:, 1::2, 1::2]
        enc_extra[l] = gcat[..., ca:]
    # g is now the gradient of the deepest encoder output (conv 9)
    for l in range(4, -1, -1):
        if l < 4:
            g = g + enc_extra[l]
        g = conv_back(k, g)
        k -= 1
        if l > 0:
            gp = conv_back(k, g)
            g = 0.25 * _up(gp)
        else:
            conv_back(k, g, want_dx=False)
        k -= 1
    return grads
