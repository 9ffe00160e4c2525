"""The serving and denoiser workloads at their configured sizes, on the GPU.

Serving: the four resident scenes at full size (the 256x256 Cornell box, the
instanced icospheres with 4,096 and with 16,384 instances, the volumes frame
with its 2048x1024 sky), and a seeded draw of requests from the serving
distribution (Poisson arrivals, resolution uniform over {256^2, 512^2,
1024x576, 1920x1080}, spp log-uniform in [16, 1024], one seed per request)
that covers every resolution and every scene. Every served film must be
bit-identical to the same request rendered alone, and one request per scene
must match the CPU oracle sample for sample on sampled 16x16 tiles.

Denoiser: trained on 256x256 renders of seed-randomised variants of the
three frame families (8-spp inputs, 1024-spp targets), then a held-out
variant at 1920x1080 and 16 spp is denoised and compared with a plain
4096-spp render of it: the denoised relative RMSE must beat the noisy one.
The same frame through sundog_render with denoise on returns the plain
render without weights and the network's output with them.
"""

import os

import numpy as np
import pytest

import sundog_amd as sd
from sundog_amd import denoise, serving
from sundog_amd.scenelib import read_pfm

pytestmark = pytest.mark.gpu


def _tile_mask(w, h, stride):
    tx = (w + 15) // 16
    m = np.zeros((h, w), bool)
    for t in range(0, tx * ((h + 15) // 16), stride):
        bx, by = (t % tx) * 16, (t // tx) * 16
        m[by:by + 16, bx:bx + 16] = True
    return m


def test_serving_workload_mix_alone_and_oracle(asset_dir, orc):
    scenes = serving.build_scenes(asset_dir)
    handles = [s.build() for s in scenes]
    srv = serving.Server(budget_bytes=64 << 30, pass_spp=16, tile=16)
    sids = [srv.add_scene(h) for h in handles]
    srv.tune("graph_passes", 4)
    reqs = serving.make_requests(7, 50.0, seed=180)
    assert {(r["width"], r["height"]) for r in reqs} == set(serving.RESOLUTIONS)
    assert {r["scene"] for r in reqs} == {0, 1, 2, 3}
    print("serving mix: %d requests, %.3g paths" % (len(reqs), sum(r["width"] * r["height"] * r["spp"]
                                                                   for r in reqs)), flush=True)
    ids = [srv.submit(sids[r["scene"]], r["width"], r["height"], r["spp"], r["seed"]) for r in reqs]
    srv.wait_idle()
    info = {d["id"]: d for d in srv.poll(64)}
    films = {}
    for rid, r in zip(ids, reqs):
        assert info[rid]["status"] == 2, info[rid]
        films[rid] = srv.fetch(rid, r["width"], r["height"])
    for rid, r in zip(ids, reqs):
        _, alone = srv.render_serial(sids[r["scene"]], r["width"], r["height"], r["spp"], r["seed"],
                                     want_image=True)
        assert np.array_equal(films[rid].view(np.uint32), alone.view(np.uint32)), r
    print("serving mix: every served film equals the request alone", flush=True)
    checked = set()
    for rid, r in zip(ids, reqs):
        if r["scene"] in checked:
            continue
        checked.add(r["scene"])
        w, h = r["width"], r["height"]
        stride = max(1, ((w + 15) // 16) * ((h + 15) // 16) // 24)
        flat = handles[r["scene"]].flatten(width=w, height=h, spp=r["spp"], seed=r["seed"])
        ref, _, _, _ = orc.render(flat, w, h, 0, r["spp"], tile_stride=stride)
        m = _tile_mask(w, h, stride)
        g, c = films[rid][m][:, :3], ref[m][:, :3]
        err = float(np.sqrt(np.sum((g - c) ** 2) / max(np.sum(c * c), 1e-30)))
        close = np.mean(np.abs(g - c) <= 1e-4 * np.maximum(1.0, np.abs(c)))
        print("scene %d %dx%d %d spp: relRMSE vs oracle on %d px %.3g, %.4f within 1e-4"
              % (r["scene"], w, h, r["spp"], m.sum(), err, close), flush=True)
        assert err < 2e-3 and close > 0.98
    assert checked == {0, 1, 2, 3}
    srv.close()


def test_denoiser_workload_1080p(asset_dir, tmp_path, monkeypatch):
    xs, ys = denoise.dataset(asset_dir, 12, 256, 256, noisy_spp=8, ref_spp=1024)
    print("denoiser: %d training renders of 256x256" % len(xs), flush=True)
    den = denoise.Denoiser(seed=1)
    losses = denoise.train(den, xs, ys, steps=1500, batch=8, crop=128, lr=1e-3)
    print("denoiser: trained, loss %.4f -> %.4f" % (np.mean(losses[:50]), np.mean(losses[-50:])), flush=True)
    w, h = 1920, 1080
    held_out = {
        "cornell": lambda spp, s: denoise.cornell_variant(1000, w, h, spp, s),
        "volumes": lambda spp, s: denoise.volumes_variant(1001, asset_dir, w, h, spp, s),
    }
    for name, fn in held_out.items():
        sess = fn(16, 5).build().session()
        sess.render_all()
        b, a, n = sess.download()
        sess.close()
        ref_sess = fn(4096, 77).build().session()
        ref_sess.render_all(chunk=256)
        ref, _, _ = ref_sess.download()
        ref_sess.close()
        out = den.infer_image(b, a, n)
        e_noisy, e_den = sd.rel_rmse(b, ref), sd.rel_rmse(out, ref)
        print("denoiser %s 1080p: relRMSE noisy %.4f denoised %.4f" % (name, e_noisy, e_den), flush=True)
        assert np.isfinite(out).all()
        assert e_den < e_noisy, (name, e_noisy, e_den)
    # on request through sundog_render: plain render without weights
    wpath = str(tmp_path / "w.sdn")
    den.save(wpath)

    def render(tag, flag):
        sc = held_out["cornell"](16, 5).build()
        path = str(tmp_path / tag)
        sc.render_file(path, denoise=flag, aov_albedo_path=path + "_a", aov_normal_path=path + "_n")
        return read_pfm(path + ".pfm"), read_pfm(path + "_a.pfm"), read_pfm(path + "_n.pfm")

    monkeypatch.delenv("SUNDOG_DENOISER_WEIGHTS", raising=False)
    plain, alb, nrm = render("plain", 0)
    fallback, _, _ = render("fallback", 1)
    assert np.array_equal(plain, fallback)
    monkeypatch.setenv("SUNDOG_DENOISER_WEIGHTS", wpath)
    dn, _, _ = render("dn", 1)
    assert not np.array_equal(dn, plain)
    assert np.allclose(dn, den.infer_image(plain, alb, nrm), rtol=1e-5, atol=1e-6)
