"""Writes this build's own GPU goldens: six scenes rendered through
sundog_render at 128x128, 64 spp, seed 7, no denoise (the reference's golden
settings, scripts/run-golden.sh:21-56, at a quarter of its 256^2 area), kept
as the 12-bit PQ / BT.2020 codes of the .pq.ppm output rounded to 8 bits.

Run on a GPU from the repository root:
    python3 tests/golden/make_gpu_goldens.py OUT_DIR
then copy OUT_DIR/gpu_goldens.npz and OUT_DIR/gpu_goldens.json here.

Before writing, every render is checked against the CPU oracle on sampled
16x16 tiles (relative RMSE and the fraction of values within 1e-4), and the
figures go into the manifest: the goldens pin this build's output over
time, the oracle pins the output itself.
"""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import sundog_amd as sd  # noqa: E402
from sundog_amd import configs  # noqa: E402
from sundog_amd.scenelib import read_pfm  # noqa: E402

SETTINGS = {"width": 128, "height": 128, "spp": 64, "seed": 7, "denoise": 0}
SCENES = ["smoke", "02-cornell-lume", "features", "cornell", "volumes", "instanced"]


def build_scene(name, asset_dir):
    """The named golden scene (fixture scenes replay the reference scene
    file's recorded builder calls; the rest are this build's configurations)."""
    if name in ("smoke", "02-cornell-lume", "features"):
        with open(os.path.join(HERE, "program_%s.json" % name)) as f:
            calls = json.load(f)["builder_calls"]
        s = sd.Scene()
        for c in calls:
            getattr(s, c["method"])(*c["args"], **c["kwargs"])
        return s
    if name == "cornell":
        return configs.cornell()
    if name == "volumes":
        return configs.volumes(asset_dir)
    if name == "instanced":
        return configs.instanced(asset_dir)
    raise KeyError(name)


def read_pq_ppm(path):
    """12-bit codes from a P6 PPM with maxval 4095 (big-endian 16-bit samples)."""
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"4095", parts[:3]
    w, h = (int(v) for v in parts[1].split())
    return np.frombuffer(parts[3], ">u2", w * h * 3).reshape(h, w, 3).astype(np.uint16)


def render_codes(scene, out_base):
    """Renders through sundog_render; returns (12-bit codes, linear film, output sha256s)."""
    scene.build().render_file(out_base, **SETTINGS)
    codes = read_pq_ppm(out_base + ".pq.ppm")
    film = read_pfm(out_base + ".pfm")
    digest = {ext: hashlib.sha256(open(out_base + ext, "rb").read()).hexdigest() for ext in (".pfm", ".pq.ppm")}
    return codes, film, digest


def to8(codes):
    return ((codes.astype(np.uint32) * 255 + 2047) // 4095).astype(np.uint8)


def psnr(codes12, golden8):
    a = codes12.astype(np.float64) / 4095.0
    b = golden8.astype(np.float64) / 255.0
    mse = float(np.mean((a - b) ** 2))
    return float("inf") if mse == 0.0 else 10.0 * np.log10(1.0 / mse)


def _oracle_check(scene, film):
    import oracle
    w, h, spp = SETTINGS["width"], SETTINGS["height"], SETTINGS["spp"]
    flat = scene.build().flatten(width=w, height=h, spp=spp, seed=SETTINGS["seed"])
    stride = 3
    ref, _, _, _ = oracle.render(flat, w, h, 0, spp, tile_stride=stride)
    tx = (w + 15) // 16
    m = np.zeros((h, w), bool)
    for t in range(0, tx * ((h + 15) // 16), stride):
        bx, by = (t % tx) * 16, (t // tx) * 16
        m[by:by + 16, bx:bx + 16] = True
    g, c = film[m][:, :3], ref[m][:, :3]
    err = float(np.sqrt(np.sum((g - c) ** 2) / max(np.sum(c * c), 1e-30)))
    close = float(np.mean(np.abs(g - c) <= 1e-4 * np.maximum(1.0, np.abs(c))))
    return err, close


def main(out_dir):
    os.makedirs(out_dir, exist_ok=True)
    asset_dir = os.path.join(out_dir, "assets")
    goldens, manifest = {}, {"settings": SETTINGS, "encoding": "PQ / BT.2020 12-bit codes of .pq.ppm, "
                             "rounded to 8 bits: (c * 255 + 2047) // 4095",
                             "metric": "PSNR over all pixels x RGB of codes normalised to [0, 1]; "
                             "tests require >= 45 dB", "scenes": {}}
    info = sd.probe()
    manifest["device"] = info.get("name", "")
    for name in SCENES:
        scene = build_scene(name, asset_dir)
        codes, film, digest = render_codes(scene, os.path.join(out_dir, name))
        err, close = _oracle_check(scene, film)
        print("%-16s oracle relRMSE %.3g, %.4f within 1e-4" % (name, err, close), flush=True)
        if not (err < 2e-3 and close > 0.98):
            raise SystemExit("%s: the render does not match the oracle; no goldens written" % name)
        goldens[name] = to8(codes)
        manifest["scenes"][name] = {"oracle_rel_rmse": err, "oracle_within_1e-4": close,
                                    "self_psnr_db": psnr(codes, goldens[name]), "sha256": digest}
    np.savez_compressed(os.path.join(out_dir, "gpu_goldens.npz"), **goldens)
    with open(os.path.join(out_dir, "gpu_goldens.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(out_dir, "gpu_goldens.npz"), flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "gpurun_out/goldens")
