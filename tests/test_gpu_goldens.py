"""Golden and determinism harness (the reference's scripts/run-golden.sh:21-65
and run-smoke.sh, re-based on this build's own GPU goldens): each golden
scene rendered through sundog_render at 128x128, 64 spp, seed 7, no denoise
must reach PSNR >= 45 dB against tests/golden/gpu_goldens.npz in normalised
PQ code-value space (the reference's metric, tests/tools/img_compare.cpp:86-93),
and the smoke scene rendered twice must give byte-identical output files.
The goldens were written by tests/golden/make_gpu_goldens.py after checking
each render against the CPU oracle (tests/golden/gpu_goldens.json)."""

import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_gpu_goldens as gg  # noqa: E402

pytestmark = pytest.mark.gpu



@pytest.fixture(scope="module")
def goldens():
    with np.load(os.path.join(HERE, "golden", "gpu_goldens.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def manifest():
    with open(os.path.join(HERE, "golden", "gpu_goldens.json")) as f:
        return json.load(f)


def test_manifest_matches_the_harness(goldens, manifest):
    assert manifest["settings"] == gg.SETTINGS
    assert sorted(goldens) == sorted(gg.SCENES)
    for name in gg.SCENES:
        assert goldens[name].shape == (gg.SETTINGS["height"], gg.SETTINGS["width"], 3)
        assert manifest["scenes"][name]["oracle_rel_rmse"] < 2e-3


@pytest.mark.parametrize("name", gg.SCENES)
def test_golden_psnr(tmp_path, asset_dir, goldens, name):
    codes, film, _ = gg.render_codes(gg.build_scene(name, asset_dir), str(tmp_path / name))
    assert np.isfinite(film).all()
    p = gg.psnr(codes, goldens[name])
    print("%s: PSNR %.2f dB against the golden" % (name, p), flush=True)
    assert p >= 45.0, (name, p)


def test_smoke_renders_are_byte_identical(tmp_path, asset_dir):
    digests = []
    for k in range(2):
        _, _, d = gg.render_codes(gg.build_scene("smoke", asset_dir), str(tmp_path / ("smoke%d" % k)))
        digests.append(d)
    assert digests[0] == digests[1]
    print("smoke sha256 %s" % digests[0][".pfm"][:16], flush=True)


def test_psnr_metric_separates_noise_from_a_changed_render(goldens):
    g = goldens["cornell"]
    codes = (g.astype(np.uint32) * 4095 // 255).astype(np.uint16)
    assert gg.psnr(codes, g) > 50.0       # the 8-bit rounding alone
    shifted = np.clip(codes.astype(np.int64) + 64, 0, 4095).astype(np.uint16)
    assert gg.psnr(shifted, g) < 45.0     # a 1.6% code shift, as from a changed exposure
    assert hashlib.sha256(g.tobytes()).hexdigest() != hashlib.sha256(g[::-1].tobytes()).hexdigest()
