"""Learned denoiser: ctypes wrapper of sundog_denoiser_*, training data from
this renderer's own output, and a training / timing CLI.

The reference hands beauty + albedo + normal films to a vendor HDR denoiser
(src/denoise.cpp:18-53, called from src/capi_render.cpp:247-256). Here the
same three films feed a BF16 U-Net whose forward and backward passes run on
the library's own MFMA kernels.

    python -m sundog_amd.denoise train --steps 2000 --out weights.sdn
    python -m sundog_amd.denoise time --height 1088 --width 1920
"""

import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import _lib
from . import configs
from .scenelib import Scene, rel_rmse, translate, scale, rotate_x, rotate_y, rotate_z


def _fp(a):
    return a.ctypes.data_as(_lib.P_f) if a is not None else None


class Denoiser(object):
    def __init__(self, seed=1, lib=None):
        self.lib = lib or _lib.load()
        self.ptr = self.lib.sundog_denoiser_create(int(seed))
        if not self.ptr:
            raise _lib.SundogError(_lib.last_error(self.lib))
        self.n_params = int(self.lib.sundog_denoiser_param_count(self.ptr))

    def close(self):
        if self.ptr:
            self.lib.sundog_denoiser_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _lib.check(self.lib.sundog_denoiser_get_params(self.ptr, _fp(p)), self.lib)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError("expected %d parameters, got %d" % (self.n_params, p.size))
        _lib.check(self.lib.sundog_denoiser_set_params(self.ptr, _fp(p)), self.lib)

    @staticmethod
    def _check_batch(inp):
        inp = np.ascontiguousarray(inp, np.float32)
        if inp.ndim == 3:
            inp = inp[None]
        if inp.ndim != 4 or inp.shape[3] != 9:
            raise ValueError("input must be [n][h][w][9]")
        return inp

    def forward_backward(self, inp, target, want_grads=True, want_pred=True):
        inp = self._check_batch(inp)
        n, h, w, _ = inp.shape
        target = np.ascontiguousarray(target, np.float32).reshape(n, h, w, 3)
        loss = ctypes.c_float(0)
        g = np.zeros(self.n_params, np.float32) if want_grads else None
        pred = np.zeros((n, h, w, 3), np.float32) if want_pred else None
        _lib.check(self.lib.sundog_denoiser_forward_backward(self.ptr, _fp(inp), _fp(target), n, h, w,
                                                             ctypes.byref(loss), _fp(g), _fp(pred)), self.lib)
        return loss.value, g, pred

    def train_step(self, inp, target, lr):
        inp = self._check_batch(inp)
        n, h, w, _ = inp.shape
        target = np.ascontiguousarray(target, np.float32).reshape(n, h, w, 3)
        loss = ctypes.c_float(0)
        _lib.check(self.lib.sundog_denoiser_train_step(self.ptr, _fp(inp), _fp(target), n, h, w, float(lr),
                                                       ctypes.byref(loss)), self.lib)
        return loss.value

    def infer(self, inp):
        inp = self._check_batch(inp)
        n, h, w, _ = inp.shape
        out = np.zeros((n, h, w, 3), np.float32)
        _lib.check(self.lib.sundog_denoiser_infer(self.ptr, _fp(inp), n, h, w, _fp(out)), self.lib)
        return out

    def infer_image(self, beauty, albedo, normal):
        """Denoise one frame of any size: pads to a multiple of 16 by edge
        replication and crops back."""
        inp = make_input(beauty, albedo, normal)
        h, w = inp.shape[:2]
        ph, pw = (-h) % 16, (-w) % 16
        if ph or pw:
            inp = np.pad(inp, ((0, ph), (0, pw), (0, 0)), mode="edge")
        return self.infer(inp)[0, :h, :w]

    def time_forward(self, n, h, w, iters=10):
        ms, fl = ctypes.c_double(0), ctypes.c_double(0)
        _lib.check(self.lib.sundog_denoiser_time_forward(self.ptr, n, h, w, iters, ctypes.byref(ms),
                                                         ctypes.byref(fl)), self.lib)
        return ms.value, fl.value

    def save(self, path):
        _lib.check(self.lib.sundog_denoiser_save(self.ptr, path.encode()), self.lib)

    def load(self, path):
        _lib.check(self.lib.sundog_denoiser_load(self.ptr, path.encode()), self.lib)


def make_input(beauty, albedo, normal):
    """Films (h, w, >=3) -> network input (h, w, 9)."""
    return np.concatenate([np.asarray(beauty, np.float32)[..., :3], np.asarray(albedo, np.float32)[..., :3],
                           np.asarray(normal, np.float32)[..., :3]], axis=-1)


# ---------------------------------------------------------------------------
# training data: seed-randomised scene variants rendered by this library
# ---------------------------------------------------------------------------

def cornell_variant(v, width, height, spp, seed):
    rng = np.random.default_rng(1000 + v)
    s = Scene()
    s.render(width=width, height=height, spp=spp, max_depth=8, seed=seed, clamp=0.0)
    cam_x, cam_y = rng.uniform(-0.4, 0.4, 2)
    s.camera(lookfrom=[cam_x, cam_y, 3.9], lookat=[0, 0, 0], vfov=float(rng.uniform(34, 46)))
    s.background_solid(color=[0, 0, 0])
    cols = rng.uniform(0.05, 0.85, (5, 3))
    s.lambert("grey", color=list(cols[0] * 0.3 + 0.5))
    s.lambert("left", color=list(cols[1]))
    s.lambert("right", color=list(cols[2]))
    s.lambert("ball", color=list(cols[3]))
    s.metal("chrome", color=list(0.5 + 0.5 * cols[4]), roughness=float(rng.uniform(0.0, 0.4)))
    s.dielectric("glass", ior=1.5, roughness=float(rng.choice([0.0, 0.0, 0.2])))
    s.emissive("lamp", color=list(rng.uniform(5, 20, 3)), intensity=1.0)
    s.add("rect", "grey", translate(0, -1, 0))
    s.add("rect", "grey", rotate_x(180), translate(0, 1, 0))
    s.add("rect", "grey", rotate_x(90), translate(0, 0, -1))
    s.add("rect", "left", rotate_z(-90), translate(-1, 0, 0))
    s.add("rect", "right", rotate_z(90), translate(1, 0, 0))
    mats = ["ball", "glass", "chrome"]
    for i in range(3):
        r = float(rng.uniform(0.2, 0.4))
        x, z = rng.uniform(-0.6, 0.6, 2)
        s.add("sphere", mats[(i + v) % 3], scale(r), translate(float(x), -1 + r, float(z)))
    lw, ld = rng.uniform(0.15, 0.4, 2)
    s.add("rect", "lamp", scale(float(lw), 1, float(ld)), rotate_x(180), translate(0, 0.999, 0))
    return s


def instanced_variant(v, asset_dir, width, height, spp, seed):
    rng = np.random.default_rng(2000 + v)
    s = configs.instanced(asset_dir, width=width, height=height, spp=spp, max_depth=8, seed=seed,
                          instances=int(rng.integers(64, 512)), subdiv=3, layout_seed=3 + v)
    return s


def volumes_variant(v, asset_dir, width, height, spp, seed):
    rng = np.random.default_rng(3000 + v)
    s = configs.volumes(asset_dir, width=width, height=height, spp=spp, max_depth=8, seed=seed, env_w=512,
                        env_h=256, flame_seed=int(rng.integers(1, 1 << 20)))
    cx, cy = rng.uniform(-0.8, 0.8), rng.uniform(-0.3, 0.5)
    s.camera(lookfrom=[cx, 1.6 + cy, 5.0], lookat=[0, 0.6, -1.0], vfov=float(rng.uniform(38, 52)))
    return s


def render_pair(scene_fn, noisy_spp, ref_spp, seed):
    """(input [h][w][9], target [h][w][3]) of one scene variant: the noisy
    render and an independent high-spp reference."""
    noisy = scene_fn(noisy_spp, seed).build()
    sess = noisy.session()
    sess.render_all()
    b, a, n = sess.download()
    sess.close()
    ref = scene_fn(ref_spp, seed + 7919).build()
    sr = ref.session()
    sr.render_all(chunk=64)
    rb, _, _ = sr.download()
    sr.close()
    return make_input(b, a, n), rb[..., :3].copy()


def dataset(asset_dir, n_variants, width, height, noisy_spp, ref_spp, first_variant=0, seed=11, verbose=False):
    xs, ys = [], []
    for v in range(first_variant, first_variant + n_variants):
        # the three benchmark frame families in turn: Cornell, volumes, instanced
        if v % 3 == 2:
            fn = lambda spp, sd, v=v: instanced_variant(v, asset_dir, width, height, spp, sd)  # noqa: E731
        elif v % 3 == 1:
            fn = lambda spp, sd, v=v: volumes_variant(v, asset_dir, width, height, spp, sd)  # noqa: E731
        else:
            fn = lambda spp, sd, v=v: cornell_variant(v, width, height, spp, sd)  # noqa: E731
        x, y = render_pair(fn, noisy_spp, ref_spp, seed + 101 * v)
        xs.append(x)
        ys.append(y)
        if verbose:
            print("dataset: variant %d (%s) %dx%d rendered" % (v, ("cornell", "volumes", "instanced")[v % 3],
                                                             width, height), flush=True)
    return np.stack(xs), np.stack(ys)


def crop_list(n, n_images, height, width, crop, seed=5):
    """A fixed training set of n crops: (image, top row, left column)."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, n_images, n), rng.integers(0, height - crop + 1, n),
                     rng.integers(0, width - crop + 1, n)], axis=1)


def random_crops(xs, ys, rng, batch, crop, picks=None):
    """A training batch: random crops, or the given (image, row, column)
    crops; each flipped horizontally at random."""
    n, h, w, _ = xs.shape
    if picks is None:
        bi = rng.integers(0, n, batch)
        oy = rng.integers(0, h - crop + 1, batch)
        ox = rng.integers(0, w - crop + 1, batch)
    else:
        bi, oy, ox = picks[:, 0], picks[:, 1], picks[:, 2]
    cx = np.stack([xs[b, y:y + crop, x:x + crop] for b, y, x in zip(bi, oy, ox)])
    cy = np.stack([ys[b, y:y + crop, x:x + crop] for b, y, x in zip(bi, oy, ox)])
    flip = rng.integers(0, 2, batch).astype(bool)
    cx[flip] = cx[flip, :, ::-1]
    cy[flip] = cy[flip, :, ::-1]
    # a horizontal flip mirrors the camera-space normal's x component
    cx[flip, :, :, 6] *= -1.0
    return cx, cy


def train(den, xs, ys, steps, batch=4, crop=64, lr=2e-3, seed=0, log_every=0, crops=None):
    """Adam steps with a decaying learning rate; with `crops` (crop_list)
    the batches walk that fixed set in a new random order every epoch."""
    rng = np.random.default_rng(seed)
    losses = []
    order = np.zeros(0, np.int64)
    for i in range(steps):
        picks = None
        if crops is not None:
            if len(order) < batch:
                order = np.concatenate([order, rng.permutation(len(crops))])
            picks, order = crops[order[:batch]], order[batch:]
        cx, cy = random_crops(xs, ys, rng, batch, crop, picks)
        cur_lr = lr * (0.1 ** (i / max(steps, 1)))
        losses.append(den.train_step(cx, cy, cur_lr))
        if log_every and (i + 1) % log_every == 0:
            print("step %d loss %.4f" % (i + 1, float(np.mean(losses[-log_every:]))), flush=True)
    return losses


def measure(args):
    """Trains on seed-randomised variants of the three frame families
    (--size squared renders, inputs at 8 and 16 spp, targets at --ref-spp,
    random --crop crops), then renders one held-out variant of each family
    at --width x --height: --input-spp noisy + denoise against a plain
    --target-spp render."""
    t0 = time.time()
    if args.noisy_spp > 0:
        xs, ys = dataset(args.asset_dir, args.variants, args.size, args.size, args.noisy_spp, args.ref_spp,
                         verbose=True)
    else:  # half the variants at 8 spp, half at 16
        xa, ya = dataset(args.asset_dir, args.variants // 2, args.size, args.size, 8, args.ref_spp, verbose=True)
        xb, yb = dataset(args.asset_dir, args.variants - args.variants // 2, args.size, args.size, 16, args.ref_spp,
                         first_variant=args.variants // 2, verbose=True)
        xs, ys = np.concatenate([xa, xb]), np.concatenate([ya, yb])
    data_s = time.time() - t0
    crops = crop_list(args.crops, len(xs), args.size, args.size, args.crop) if args.crops > 0 else None
    den = Denoiser(seed=1)
    t1 = time.time()
    losses = train(den, xs, ys, args.steps, batch=8, crop=args.crop, lr=1e-3, log_every=max(args.steps // 10, 1),
                   crops=crops)
    train_s = time.time() - t1
    if args.out:
        den.save(args.out)
    w, h = args.width, args.height
    held_out = {
        "cornell": lambda spp, sd: cornell_variant(1000, w, h, spp, sd),
        "volumes": lambda spp, sd: volumes_variant(1001, args.asset_dir, w, h, spp, sd),
        "instanced": lambda spp, sd: instanced_variant(1002, args.asset_dir, w, h, spp, sd),
    }
    frames = {}
    ref_s = 0.0
    for name, fn in held_out.items():
        s = fn(args.input_spp, 5).build().session()
        s.render_all()
        b, a, n = s.download()
        s.close()
        t2 = time.time()
        r = fn(args.target_spp, 77).build().session()
        r.render_all(chunk=256)
        ref, _, _ = r.download()
        r.close()
        ref_s += time.time() - t2
        out = den.infer_image(b, a, n)
        frames[name] = {"rel_rmse_noisy": rel_rmse(b, ref), "rel_rmse_denoised": rel_rmse(out, ref)}
        print("%s: noisy %.4f denoised %.4f" % (name, frames[name]["rel_rmse_noisy"],
                                                 frames[name]["rel_rmse_denoised"]), flush=True)
    ms, fl = den.time_forward(1, (h + 15) // 16 * 16, (w + 15) // 16 * 16, iters=5)
    res = {"variants": args.variants, "render_size": args.size, "noisy_spp": args.noisy_spp or "8 and 16",
           "ref_spp": args.ref_spp, "crops": args.crops or "random", "crop": args.crop,
           "train_steps": args.steps, "train_s": train_s, "data_s": data_s,
           "final_loss": float(np.mean(losses[-50:])),
           "frame": [w, h], "input_spp": args.input_spp, "target_spp": args.target_spp, "target_render_s": ref_s,
           "held_out": frames,
           "denoise_forward_ms": ms, "denoise_gflop": fl / 1e9,
           "mfma_frac_bf16_dense": fl / (ms * 1e-3) / 2.5e15}
    print(json.dumps(res), flush=True)
    return res


def train_steps(args):
    """Training steps (forward, backward, Adam) on synthetic crops: log-HDR
    inputs and targets with per-pixel noise; reports ms per step and the
    step's FLOPs (forward + backward-data + backward-weights = 3x forward)."""
    rng = np.random.default_rng(3)
    n, c = args.batch, args.crop
    base = rng.gamma(2.0, 0.5, (n, c, c, 3)).astype(np.float32)
    inp = np.concatenate([base * rng.gamma(4.0, 0.25, (n, c, c, 3)), rng.uniform(0, 1, (n, c, c, 3)),
                          rng.normal(0, 0.5, (n, c, c, 3))], axis=-1).astype(np.float32)
    den = Denoiser(seed=1)
    for _ in range(2):
        den.train_step(inp, base, 1e-3)
    _, fwd_flop = den.time_forward(n, c, c, iters=1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        den.train_step(inp, base, 1e-3)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    flop = 3.0 * fwd_flop
    out = {"train_step_ms": ms, "batch": n, "crop": c, "step_gflop": flop / 1e9,
           "tflops": flop / (ms * 1e-3) / 1e12, "mfma_frac_bf16_dense": flop / (ms * 1e-3) / 2.5e15}
    print(json.dumps(out), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--steps", type=int, default=2000)
    t.add_argument("--variants", type=int, default=12)
    t.add_argument("--size", type=int, default=128)
    t.add_argument("--noisy-spp", type=int, default=8)
    t.add_argument("--ref-spp", type=int, default=1024)
    t.add_argument("--batch", type=int, default=8)
    t.add_argument("--crop", type=int, default=64)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", default="")
    t.add_argument("--asset-dir", default=os.path.join(tempfile.gettempdir(), "sundog_assets"))
    ms = sub.add_parser("measure", help="train, then denoise a held-out 1080p frame and compare with a "
                                        "plain high-spp render of it")
    ms.add_argument("--steps", type=int, default=7500)
    ms.add_argument("--variants", type=int, default=36)
    ms.add_argument("--size", type=int, default=1024)
    ms.add_argument("--ref-spp", type=int, default=4096)
    ms.add_argument("--noisy-spp", type=int, default=8, help="input spp of the training renders (0: half 8, half 16)")
    ms.add_argument("--crops", type=int, default=20000,
                    help="a fixed set of this many training crops (0: fresh random crops every step)")
    ms.add_argument("--crop", type=int, default=256)
    ms.add_argument("--width", type=int, default=1920)
    ms.add_argument("--height", type=int, default=1080)
    ms.add_argument("--input-spp", type=int, default=16)
    ms.add_argument("--target-spp", type=int, default=4096)
    ms.add_argument("--out", default="")
    ms.add_argument("--asset-dir", default=os.path.join(tempfile.gettempdir(), "sundog_assets"))
    tm = sub.add_parser("time")
    tm.add_argument("--batch", type=int, default=1)
    tm.add_argument("--height", type=int, default=1088)
    tm.add_argument("--width", type=int, default=1920)
    tm.add_argument("--iters", type=int, default=10)
    ts = sub.add_parser("train-steps", help="time training steps on synthetic crops (a short program to profile)")
    ts.add_argument("--steps", type=int, default=20)
    ts.add_argument("--batch", type=int, default=8)
    ts.add_argument("--crop", type=int, default=128)
    args = ap.parse_args(argv)

    if args.cmd == "train-steps":
        return train_steps(args)

    if args.cmd == "time":
        den = Denoiser(seed=1)
        ms, fl = den.time_forward(args.batch, args.height, args.width, args.iters)
        out = {"forward_ms": ms, "gflop": fl / 1e9, "tflops": fl / (ms * 1e-3) / 1e12,
               "mfma_frac_bf16_dense": fl / (ms * 1e-3) / 2.5e15,
               "shape": [args.batch, args.height, args.width]}
        print(json.dumps(out), flush=True)
        return out

    if args.cmd == "measure":
        return measure(args)

    t0 = time.time()
    xs, ys = dataset(args.asset_dir, args.variants, args.size, args.size, args.noisy_spp, args.ref_spp)
    xt, yt = dataset(args.asset_dir, 3, args.size, args.size, args.noisy_spp, args.ref_spp,
                     first_variant=args.variants)
    print("data: %d train + %d held-out images, %.1f s" % (len(xs), len(xt), time.time() - t0), flush=True)
    den = Denoiser(seed=1)
    t1 = time.time()
    losses = train(den, xs, ys, args.steps, args.batch, args.crop, args.lr, log_every=max(args.steps // 10, 1))
    train_s = time.time() - t1
    noisy = [rel_rmse(x[..., :3], y) for x, y in zip(xt, yt)]
    den_out = den.infer(xt)
    denoised = [rel_rmse(d, y) for d, y in zip(den_out, yt)]
    out = {"steps": args.steps, "train_s": train_s, "final_loss": float(np.mean(losses[-50:])),
           "heldout_rel_rmse_noisy": noisy, "heldout_rel_rmse_denoised": denoised}
    print(json.dumps(out), flush=True)
    if args.out:
        den.save(args.out)
    return out


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
