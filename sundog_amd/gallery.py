"""Reference gallery scenes rendered the way the reference benchmarks them.

The reference's benchmark (scripts/run-benchmark.sh:50-69, published in
docs/BENCHMARKS.md) renders each gallery scene at 1920x1080 / 64 spp
without denoising and reports the render loop's wall time and Mrays/s from
the render's stats JSON (src/capi_render.cpp:230-292: radiance plus traced
shadow queries over the spp loop's wall time). This module replays a
scene's recorded backend call program (the reference scenelib's output for
that scene file, a JSON fixture) through this library's C ABI and renders
it with `sundog_render` under the same overrides, so the figures are the
same quantities as the published ones.

Only gallery scenes that need no asset files are recorded: 01-marble-run
(quadrics, grid texture) and 02-cornell-lume (quadrics, area lights).

    python -m sundog_amd.gallery [--size 1920x1080] [--spp 64] PROGRAM.json ...
"""

import argparse
import json
import os
import sys
import tempfile

from . import _lib
from .scenelib import SceneHandle, apply_program

# docs/BENCHMARKS.md:9-10 (NVIDIA GB200, OptiX with software BVH traversal)
PUBLISHED = {
    "01-marble-run": {"render_s": 0.254, "mrays_per_s": 1209.0, "gpu": "GB200 (software traversal)",
                      "source": "docs/BENCHMARKS.md:9"},
    "02-cornell-lume": {"render_s": 1.02, "mrays_per_s": 819.0, "gpu": "GB200 (software traversal)",
                        "source": "docs/BENCHMARKS.md:10"},
}

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE_DIR = os.path.join(os.path.dirname(HERE), "tests", "golden")


def program_path(name):
    return os.path.join(FIXTURE_DIR, "program_%s.json" % name)


def load_program(path):
    with open(path) as f:
        return json.load(f)["program"]


def scene_name(path):
    base = os.path.basename(path)
    return base[len("program_"):-len(".json")] if base.startswith("program_") else base


def measure(path, width=1920, height=1080, spp=64, lib=None, keep_dir=None):
    """Renders the recorded program at `path` through `sundog_render` and
    returns the stats the reference's benchmark reads, plus the published
    figures for the scene where there are any."""
    lib = lib or _lib.load()
    name = scene_name(path)
    h = SceneHandle(apply_program(load_program(path), lib), lib)
    try:
        with tempfile.TemporaryDirectory() as tmp:
            d = keep_dir or tmp
            stats_path = os.path.join(d, "%s.stats.json" % name)
            h.render_file(os.path.join(d, "%s.pfm" % name), width=width, height=height, spp=spp, denoise=0,
                          stats_path=stats_path, scene_name=name + ".py")
            with open(stats_path) as f:
                st = json.load(f)
    finally:
        h.close()
    out = {"scene": name, "width": st["width"], "height": st["height"], "spp": st["spp"],
           "max_depth": st["max_depth"], "render_s": st["timings_ms"]["render"] / 1000.0,
           "mrays_per_s": st["mrays_per_sec"], "rays_traced": st["rays_traced"],
           "paths_per_s": st["width"] * st["height"] * st["spp"] / (st["timings_ms"]["render"] / 1000.0),
           "objects": st["scene_stats"]["objects"], "lights": st["scene_stats"]["lights"],
           "peak_vram_mb": st["peak_vram_mb"]}
    pub = PUBLISHED.get(name)
    if pub is not None and (width, height, spp) == (1920, 1080, 64):
        out["published"] = dict(pub)
        out["mrays_ratio_vs_published"] = out["mrays_per_s"] / pub["mrays_per_s"]
        out["render_time_ratio_vs_published"] = pub["render_s"] / out["render_s"]
    return out


def read_pq_ppm(path):
    """The 12-bit PQ / BT.2020 code values `sundog_render` writes beside its
    PFM (`<out>.pq.ppm`, 16-bit big-endian, maxval 4095), as (h, w, 3)."""
    import numpy as np
    with open(path, "rb") as f:
        data = f.read()
    head, pos = [], 0
    while len(head) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        head.append(data[pos:end])
        pos = end
    pos += 1
    if head[0] != b"P6" or int(head[3]) != 4095:
        raise ValueError("%s: not a 12-bit PQ PPM" % path)
    w, h = int(head[1]), int(head[2])
    return np.frombuffer(data, ">u2", w * h * 3, pos).reshape(h, w, 3).astype(np.uint16)


def psnr_codes(a, b, peak=4095.0):
    """PSNR between two code-value images, as the reference's img_compare
    computes it (tests/tools/img_compare.cpp:86-93): mean over every RGB
    sample of ((a - b) / peak)^2, 10 log10(1 / mse)."""
    import numpy as np
    d = (a.astype(np.float64) - b.astype(np.float64)) / peak
    mse = float(np.mean(d * d))
    return float("inf") if mse <= 0.0 else 10.0 * np.log10(1.0 / mse)


def render_codes(path, width, height, spp, seed, workdir, lib=None):
    """Renders the recorded program at `path` through `sundog_render`
    (no denoising) and returns its 12-bit PQ code values."""
    lib = lib or _lib.load()
    h = SceneHandle(apply_program(load_program(path), lib), lib)
    try:
        out = os.path.join(workdir, "%s_%d.pfm" % (scene_name(path), spp))
        h.render_file(out, width=width, height=height, spp=spp, seed=seed, denoise=0)
    finally:
        h.close()
    return read_pq_ppm(out[:-4] + ".pq.ppm")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("programs", nargs="*", help="program_<scene>.json files (default: the recorded gallery scenes)")
    ap.add_argument("--size", default="1920x1080")
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--repeat", type=int, default=1)
    args = ap.parse_args(argv)
    w, h = (int(v) for v in args.size.lower().split("x"))
    paths = args.programs or [program_path(n) for n in sorted(PUBLISHED)]
    for p in paths:
        for _ in range(args.repeat):
            print(json.dumps(measure(p, w, h, args.spp)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
