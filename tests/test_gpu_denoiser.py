"""GPU denoiser: MFMA forward/backward against the CPU restatement (with the
same BF16 rounding points), determinism, training, and the gain of a
network trained on this renderer's output over the noisy input."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import denoiser_ref as dr  # noqa: E402

from sundog_amd import denoise  # noqa: E402

pytestmark = pytest.mark.gpu


def _batch(rng, n, h, w):
    x = np.zeros((n, h, w, 9), np.float32)
    x[..., :3] = rng.gamma(0.8, 1.5, (n, h, w, 3))
    x[..., 3:6] = rng.uniform(0, 1, (n, h, w, 3))
    nn = rng.standard_normal((n, h, w, 3))
    x[..., 6:9] = nn / np.linalg.norm(nn, axis=-1, keepdims=True)
    y = (x[..., :3] * rng.uniform(0.5, 1.5, (n, h, w, 3))).astype(np.float32)
    return x, y


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# (1, 48, 80): levels of 40x24, 20x12, 10x6 and 5x3 pixels, partial 16x8
# tiles of the spatially tiled convolution in both directions
@pytest.mark.parametrize("shape", [(2, 32, 32), (1, 32, 48), (1, 48, 80)])
def test_forward_backward_matches_oracle(shape):
    n, h, w = shape
    den = denoise.Denoiser(seed=3)
    p = den.get_params()
    rng = np.random.default_rng(h * 100 + w)
    # non-zero biases so every bias path is exercised
    offs, _ = dr.param_offsets()
    for (cin, cout, _), (wo, bo) in zip(dr.conv_specs(), offs):
        p[bo:bo + cout] = 0.02 * rng.standard_normal(cout)
    den.set_params(p)
    x, y = _batch(rng, n, h, w)
    loss, grads, pred = den.forward_backward(x, y)
    pred_o, cache = dr.forward(p, x, emulate=True)
    fin_o = dr.final_output(pred_o, x, emulate=True)
    loss_o, gpred = dr.loss_grad(fin_o, y, emulate=True)
    grads_o = dr.backward(p, cache, gpred, emulate=True)
    # forward: FP32 accumulation order differs from the restatement's, and a
    # BF16 store can then round the other way; both stay far below a layout
    # or indexing error, which gives O(1) differences
    net = pred - np.log1p(np.maximum(x[..., :3], 0.0))   # the network's own output
    assert _rel(net, pred_o[..., :3]) < 1e-2, _rel(net, pred_o[..., :3])
    assert _rel(pred, fin_o) < 1e-3
    assert abs(loss - loss_o) <= 1e-3 * abs(loss_o)
    # gradients: a BF16 rounding step that lands a pre-activation on the
    # other side of zero flips that pixel's ReLU mask, an O(1) change at one
    # pixel; at the 6x10 and 3x5 levels of the 48x80 input a few such flips
    # give relative errors up to 3.4% in the deepest convolutions (the same
    # with the tap-major kernel for every layer), against O(1) for a layout
    # or indexing error
    for k, ((cin, cout, _), (wo, bo)) in enumerate(zip(dr.conv_specs(), offs)):
        ew = _rel(grads[wo:bo], grads_o[wo:bo])
        eb = _rel(grads[bo:bo + cout], grads_o[bo:bo + cout])
        assert ew < 5e-2 and eb < 5e-2, (k, ew, eb)


def test_inference_forward_equals_training_forward():
    """The inference forward reads the pooled and concatenated layer inputs
    from their sources and writes the final image from the output layer;
    the training forward materialises them. Same operations, so the final
    log-space image is the same (compared through expm1, which the inference
    entry point applies on the GPU)."""
    den = denoise.Denoiser(seed=7)
    rng = np.random.default_rng(11)
    for n, h, w in [(1, 48, 80), (2, 32, 32)]:
        x, y = _batch(rng, n, h, w)
        _, _, fin = den.forward_backward(x, y, want_grads=False)
        out = den.infer(x)
        ref = np.expm1(fin.astype(np.float64))
        err = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-6)
        assert err.max() < 2e-6, float(err.max())


def test_deterministic_and_params_unchanged():
    den = denoise.Denoiser(seed=5)
    p0 = den.get_params()
    rng = np.random.default_rng(2)
    x, y = _batch(rng, 2, 32, 32)
    l1, g1, p1 = den.forward_backward(x, y)
    l2, g2, p2 = den.forward_backward(x, y)
    assert l1 == l2
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32))
    assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))
    assert np.array_equal(den.get_params(), p0)


def test_save_load_and_infer(tmp_path):
    den = denoise.Denoiser(seed=9)
    rng = np.random.default_rng(4)
    x, y = _batch(rng, 1, 32, 32)
    for _ in range(3):
        den.train_step(x, y, 1e-3)
    out = den.infer(x)
    path = str(tmp_path / "w.sdn")
    den.save(path)
    d2 = denoise.Denoiser(seed=1)
    d2.load(path)
    assert np.array_equal(d2.get_params(), den.get_params())
    assert np.array_equal(d2.infer(x), out)
    assert out.shape == (1, 32, 32, 3) and np.isfinite(out).all()
    # infer returns expm1 of the final log-space output
    _, _, pred = den.forward_backward(x, y)
    assert np.allclose(out, np.expm1(pred), rtol=1e-5, atol=1e-6)
    # an untrained network with a zero output layer is the identity
    p = den.get_params()
    offs, _ = dr.param_offsets()
    p[offs[-1][0]:] = 0.0
    den.set_params(p)
    ident = den.infer(x)
    assert np.allclose(ident, np.maximum(x[..., :3], 0.0), rtol=2e-6, atol=1e-6)


def _box3(img):
    n, h, w, c = img.shape
    p = np.pad(img, ((0, 0), (1, 1), (1, 1), (0, 0)), mode="edge")
    return sum(p[:, dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)) / 9.0


def test_training_reduces_loss():
    """A learnable map (a 3x3 box filter of the beauty) is fitted on a
    fixed batch: the loss must fall to under half its starting value."""
    den = denoise.Denoiser(seed=1)
    rng = np.random.default_rng(8)
    x, _ = _batch(rng, 4, 32, 32)
    y = _box3(x[..., :3]).astype(np.float32)
    first = den.train_step(x, y, 1e-3)
    for _ in range(150):
        last = den.train_step(x, y, 1e-3)
    assert last < 0.5 * first, (first, last)


def test_denoiser_gain_on_renders(asset_dir):
    """Trained on noisy/reference pairs of seed-randomised scene variants,
    the network must lower the error of held-out noisy renders."""
    xs, ys = denoise.dataset(asset_dir, 18, 128, 128, noisy_spp=8, ref_spp=1024)
    xt, yt = denoise.dataset(asset_dir, 3, 128, 128, noisy_spp=8, ref_spp=1024, first_variant=40)
    den = denoise.Denoiser(seed=1)
    denoise.train(den, xs, ys, steps=2000, batch=8, crop=64, lr=1e-3)
    out = den.infer(xt)
    noisy = np.mean([denoise.rel_rmse(x[..., :3], y) for x, y in zip(xt, yt)])
    den_err = np.mean([denoise.rel_rmse(d, y) for d, y in zip(out, yt)])
    print("held-out relRMSE noisy %.4f denoised %.4f" % (noisy, den_err))
    assert den_err < 0.8 * noisy, (noisy, den_err)


def test_render_denoise_on_request(asset_dir, tmp_path, monkeypatch):
    """sundog_render with denoise=1: the trained network's output when
    SUNDOG_DENOISER_WEIGHTS names weights, the plain render otherwise."""
    from sundog_amd import configs
    from sundog_amd.scenelib import read_pfm
    den = denoise.Denoiser(seed=2)
    rng = np.random.default_rng(6)
    x, y = _batch(rng, 2, 32, 32)
    for _ in range(5):
        den.train_step(x, y, 1e-3)
    wpath = str(tmp_path / "w.sdn")
    den.save(wpath)

    def render(name, denoise_flag):
        sc = configs.cornell(width=40, height=24, spp=8).build()
        path = str(tmp_path / name)
        sc.render_file(path, denoise=denoise_flag, aov_albedo_path=path + "_a", aov_normal_path=path + "_n")
        return read_pfm(path + ".pfm"), read_pfm(path + "_a.pfm"), read_pfm(path + "_n.pfm")

    monkeypatch.delenv("SUNDOG_DENOISER_WEIGHTS", raising=False)
    plain, alb, nrm = render("plain", 0)
    fallback, _, _ = render("fallback", 1)
    assert np.array_equal(plain, fallback)
    monkeypatch.setenv("SUNDOG_DENOISER_WEIGHTS", wpath)
    dn, _, _ = render("dn", 1)
    ref = den.infer_image(plain, alb, nrm)
    assert not np.array_equal(dn, plain)
    assert np.allclose(dn, ref, rtol=1e-5, atol=1e-6)
