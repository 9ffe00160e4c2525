"""The scene-file command line (the reference's scenes/scenelib.py:898-929)
and the smoke checks of scripts/run-smoke.sh:37-59: --probe reports a GPU,
renders at --size / --spp (plain and --denoise) leave output files, and
--stats writes JSON with the timings_ms / rays_traced / device keys. Flag
parsing runs on the CPU; the renders run a generated scene file that does
`from scenelib import ...` through `python -m sundog_amd.scenelib`, as a
reference scene file is run with this library."""

import json
import math
import os
import subprocess
import sys

import pytest

from sundog_amd import scenelib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCENE_FILE = '''from scenelib import Scene, scale, translate

s = Scene()
s.render(width=256, height=256, spp=16, max_depth=8, seed=7)
s.camera(lookfrom=[0, 1.5, 5], lookat=[0, 0.7, 0], vfov=35)
s.background_gradient(horizon=[1.0, 1.0, 1.0], zenith=[0.4, 0.6, 1.0])
s.texture("floor", "checker", a=[0.85, 0.85, 0.85], b=[0.15, 0.15, 0.15], scale=[8, 8])
s.lambert("ground", texture="floor")
s.lambert("ball", color=[0.7, 0.3, 0.3])
s.point_light(position=[3, 4, 2], intensity=[40, 40, 40], radius=0.4)
s.add("rect", "ground", scale(4))
s.add("sphere", "ball", scale(0.7), translate(0, 0.7, 0))

if __name__ == "__main__":
    s.run(out="cli.avif")
'''


def test_flags_map_to_render_options():
    o = scenelib.parse_args(["--size", "64x48", "--spp", "4", "--seed", "9", "--denoise", "--opaque-shadows",
                             "--clamp", "2.5", "--stats", "s.json", "--quiet"], "x.avif")
    assert (o.out, o.width, o.height, o.spp, o.seed) == ("x.avif", 64, 48, 4, 9)
    r = scenelib.render_overrides(o, "scene.py")
    assert r["denoise"] == 1 and r["transparent_shadows"] == 0 and r["clamp"] == 2.5
    assert r["stats_path"] == "s.json" and r["quiet"] == 1 and r["scene_name"] == "scene.py"
    d = scenelib.render_overrides(scenelib.parse_args([], "y.avif"), "s")
    assert (d["spp"], d["width"], d["height"], d["seed"], d["denoise"], d["transparent_shadows"]) == (-1,) * 6
    assert math.isnan(d["clamp"]) and math.isnan(d["physics_time"]) and "stats_path" not in d
    assert scenelib.parse_args(["--no-denoise"], "z").denoise == 0
    with pytest.raises(SystemExit):
        scenelib.parse_args(["--size", "64"], "z")


def _run(tmp_path, *flags):
    scene = tmp_path / "cli_scene.py"
    scene.write_text(SCENE_FILE)
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    env.pop("SUNDOG_DENOISER_WEIGHTS", None)
    return subprocess.run([sys.executable, "-m", "sundog_amd.scenelib", str(scene)] + list(flags), env=env,
                          cwd=str(tmp_path), stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=120)


@pytest.mark.gpu
def test_smoke_checks_through_the_scene_cli(tmp_path):
    r = _run(tmp_path, "--probe")
    assert r.returncode == 0 and r.stdout.startswith("GPU:"), r.stderr
    print(r.stdout, flush=True)
    for name, extra in (("plain", []), ("dn", ["--denoise"])):
        out = str(tmp_path / ("smoke-%s.avif" % name))
        r = _run(tmp_path, "--out", out, "--size", "64x64", "--spp", "4", "--quiet", *extra)
        assert r.returncode == 0, r.stderr
        for path in (out + ".pfm",):
            assert os.path.getsize(path) > 1024, path
    out = str(tmp_path / "smoke-st.avif")
    stats = str(tmp_path / "smoke.stats.json")
    r = _run(tmp_path, "--out", out, "--size", "64x64", "--spp", "4", "--quiet", "--stats", stats)
    assert r.returncode == 0, r.stderr
    with open(stats) as f:
        d = json.load(f)
    for key in ("timings_ms", "rays_traced", "device"):
        assert key in d, key
    assert d["timings_ms"]["render"] > 0 and d["rays_traced"] > 64 * 64 * 4
    assert (d["width"], d["height"], d["spp"]) == (64, 64, 4)
    print("stats: render %.3f ms, %d rays" % (d["timings_ms"]["render"], d["rays_traced"]), flush=True)
    bad = _run(tmp_path, "--size", "64")
    assert bad.returncode != 0
