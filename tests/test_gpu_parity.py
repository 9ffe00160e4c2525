"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact: PCG32 streams, closest-hit (t, instance, primitive) from the BVH
against brute force on both sides, environment-CDF bins. Within tolerance:
shadow transmittance (unordered float products), and images (relative RMSE
over pixels x RGB of the linear film; the renderers share PCG streams, so
they only part where a float difference flips a branch).
"""

import os

import numpy as np
import pytest

import sundog_amd as sd
from sundog_amd import configs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rays_in_box(n, seed, lo=-0.95, hi=0.95):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = 1e16
    return r


def _rays_looking_at(n, seed, eye_lo, eye_hi, target_lo, target_hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(eye_lo, eye_hi, (n, 3))
    t = rng.uniform(target_lo, target_hi, (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = 1e16
    return r


def test_device_pcg32_matches_oracle_and_published_vectors(orc):
    dev = sd.device_pcg32(42, 54, 6)
    ref = orc.pcg32(42, 54, 6)
    gold = np.load(os.path.join(GOLDEN, "pcg32_demo.npz"))["round1"]
    assert np.array_equal(dev, gold)
    assert np.array_equal(ref, gold)
    # a path stream key as the renderer uses it: (pixel << 32) ^ sample
    key = (1234 << 32) ^ 7
    assert np.array_equal(sd.device_pcg32(key, 1, 64), orc.pcg32(key, 1, 64))


@pytest.mark.parametrize("scene_name", ["cornell", "instanced_mini"])
def test_closest_hit_bit_exact(scene_name, orc, asset_dir):
    if scene_name == "cornell":
        sc = configs.cornell(width=32, height=32, spp=1).build()
        rays = _rays_in_box(4096, 11)
    else:
        sc = configs.instanced(asset_dir, width=32, height=18, spp=1, instances=24, subdiv=2).build()
        rays = _rays_looking_at(4096, 12, [-60, 2, -60], [60, 30, 60], [-50, 0, -50], [50, 3, 50])
    sess = sc.session()
    flat = sc.flatten()
    t_bvh, i_bvh, p_bvh = sess.trace(rays)
    t_bf, i_bf, p_bf = sess.trace(rays, brute=True)
    t_c, i_c, p_c = orc.trace(flat, rays, brute=True)
    t_cb, i_cb, p_cb = orc.trace(flat, rays, brute=False)
    hits = int((i_bvh >= 0).sum())
    assert hits > len(rays) // 4, "too few hits to be a meaningful test (%d)" % hits
    for (t, i, p) in ((t_bf, i_bf, p_bf), (t_c, i_c, p_c), (t_cb, i_cb, p_cb)):
        assert np.array_equal(i_bvh, i)
        assert np.array_equal(p_bvh, p)
        assert np.array_equal(t_bvh.view(np.uint32), t.view(np.uint32))


def test_non_finite_rays_miss(orc, asset_dir):
    """Rays with a NaN or infinite origin or direction hit nothing: closest
    hit = miss and shadow = unoccluded, on the GPU BVH, GPU brute force and
    the oracle alike. (NaN slabs pass every box test, so such a query visits
    every node; no primitive test accepts it. The scene is small.)"""
    sc = configs.instanced(asset_dir, width=32, height=18, spp=1, instances=24, subdiv=2).build()
    sess = sc.session()
    flat = sc.flatten()
    rays = _rays_looking_at(64, 21, [-60, 2, -60], [60, 30, 60], [-50, 0, -50], [50, 3, 50])
    bad = rays.copy()
    bad[0:16, 4] = np.nan            # NaN direction component
    bad[16:32, 0:3] = np.nan         # NaN origin
    bad[32:48, 1] = np.inf           # infinite origin component
    bad[48:64, 4:7] = [np.inf, 0.0, 0.0]
    t, i, p = sess.trace(bad)
    t_bf, i_bf, _ = sess.trace(bad, brute=True)
    t_c, i_c, _ = orc.trace(flat, bad, brute=True)
    assert (i == -1).all() and (p == -1).all() and np.isinf(t).all()
    assert (i_bf == -1).all() and (i_c == -1).all()
    # the finite rays of the same batch still hit
    t2, i2, _ = sess.trace(rays)
    assert (i2 >= 0).sum() > 16
    bad[:, 7] = 50.0
    assert np.array_equal(sess.shadow(bad), np.ones((64, 3), np.float32))


def test_shadow_transmittance_matches_oracle(orc, asset_dir):
    sc = configs.volumes(asset_dir, width=32, height=18, spp=1, env_w=64, env_h=32).build()
    sess = sc.session()
    flat = sc.flatten()
    rays = _rays_looking_at(2048, 13, [-3, 0.2, 1], [3, 2.5, 3], [-3, -0.5, -3], [3, 1.5, -1.5])
    seg = np.linalg.norm(np.array([0, 0, -4.0]))
    rays[:, 7] = seg
    tr_g = sess.shadow(rays)
    tr_c = orc.shadow(flat, rays)
    frac = ((tr_c > 0) & (tr_c < 1)).any(axis=1).mean()
    assert frac > 0.05, "no ray crossed a transmissive surface"
    np.testing.assert_allclose(tr_g, tr_c, rtol=2e-5, atol=1e-6)


def test_flame_field_matches_oracle(orc, asset_dir):
    """The device flame field stops evaluating noise octaves once the density
    mask is provably zero; the oracle evaluates every octave. On a dense grid
    over the flame's box plus random points: wherever the device returns a
    zero field the oracle's is zero too (up to a value rounding can leave at
    the mask's edge), and elsewhere sigma and emission agree to FP32
    contraction differences."""
    sc = configs.volumes(asset_dir, width=32, height=18, spp=1, env_w=64, env_h=32).build()
    sess = sc.session()
    flat = sc.flatten()
    # the volumes flame: base (-1.8, 0.06, -1.2), height 2, radius 0.5, box bounds
    gx, gy, gz = np.meshgrid(np.linspace(-2.35, -1.25, 45), np.linspace(0.0, 2.1, 85),
                             np.linspace(-1.75, -0.65, 45), indexing="ij")
    grid = np.stack([gx, gy, gz], -1).reshape(-1, 3)
    rnd = np.random.default_rng(5).uniform([-2.3, 0.06, -1.7], [-1.3, 2.06, -0.7], (50000, 3))
    p = np.concatenate([grid, rnd]).astype(np.float32)
    sg_g, em_g = sess.flame_field(0, p)
    sg_c, em_c = orc.flame_field(flat, 0, p)
    zero_g = sg_g == 0
    assert 0.2 < zero_g.mean() < 0.95, zero_g.mean()        # both regimes sampled
    assert np.all(sg_c[zero_g] <= 1e-5 * sg_c.max()), float(sg_c[zero_g].max())
    assert np.all(em_c[zero_g] <= 1e-5 * em_c.max())
    np.testing.assert_allclose(sg_g, sg_c, rtol=1e-4, atol=1e-5 * sg_c.max())
    np.testing.assert_allclose(em_g, em_c, rtol=1e-4, atol=1e-5 * em_c.max())


def test_env_bins_bit_exact(orc, asset_dir):
    sc = configs.volumes(asset_dir, width=32, height=18, spp=1, env_w=256, env_h=128).build()
    sess = sc.session()
    flat = sc.flatten()
    u = np.random.default_rng(5).random((20000, 2)).astype(np.float32)
    r_g, c_g = sess.env_bins(u)
    r_c, c_c = orc.env_bins(flat, u)
    assert np.array_equal(r_g, r_c)
    assert np.array_equal(c_g, c_c)


def test_env_bins_search_modes(orc, asset_dir):
    """The full 2048x1024 sky (its CDFs have 1,025 and 2,049 entries, and a
    0.5 degree sun makes them nearly flat around one steep step): the
    renderer's per-lane 16-ary table and the wave-cooperative search give
    the oracle's bisection bins for 2^20 samples, including u on CDF entries
    exactly and u in the flat stretches; kernel times of both are printed."""
    sc = configs.volumes(asset_dir, width=32, height=18, spp=1).build()
    sess = sc.session()
    flat = sc.flatten()
    rng = np.random.default_rng(9)
    u = rng.random((1 << 20, 2)).astype(np.float32)
    # samples on the CDFs' own entries (ties: the largest index wins)
    marg, cond = orc.env_cdf(flat, 2048, 1024)
    below1 = np.nextafter(np.float32(1.0), np.float32(0.0))
    u[:1024, 0] = np.minimum(marg[:1024], below1)
    u[1024:3072, 1] = np.minimum(cond[512, :2048], below1)
    u[1024:3072, 0] = marg[512]      # rows near the horizon
    u[-1] = [0.0, 0.0]
    u[-2] = [np.nextafter(np.float32(1.0), np.float32(0.0))] * 2
    r_c, c_c = orc.env_bins(flat, u)
    times = {}
    for mode in (0, 1):
        r_g, c_g, ms = sess.env_bins_mode(u, mode)
        assert np.array_equal(r_g, r_c), mode
        assert np.array_equal(c_g, c_c), mode
        times[mode] = ms
    print("env bins for %d samples: per-lane table %.3f ms, wave-cooperative %.3f ms"
          % (len(u), times[0], times[1]), flush=True)


def _render_pair(scene, w, h, spp, orc, tile_stride=1):
    sess = scene.session()
    sess.render_all()
    gpu, alb, nrm = sess.download()
    flat = scene.flatten()
    ref, ralb, rnrm, _ = orc.render(flat, w, h, 0, spp, tile_stride=tile_stride)
    return sess, gpu, ref


def _tile_mask(w, h, stride):
    tx = (w + 15) // 16
    m = np.zeros((h, w), bool)
    for t in range(0, tx * ((h + 15) // 16), stride):
        bx, by = (t % tx) * 16, (t // tx) * 16
        m[by:by + 16, bx:bx + 16] = True
    return m


def test_cornell_image_parity(orc):
    w = h = 64
    spp = 256
    sc = configs.cornell(width=w, height=h, spp=spp).build()
    _, gpu, ref = _render_pair(sc, w, h, spp, orc)
    assert np.isfinite(gpu).all()
    err = sd.rel_rmse(gpu, ref)
    assert err < 0.02, err


def test_repeat_render_bit_identical():
    sc = configs.cornell(width=48, height=40, spp=32).build()
    s1 = sc.session()
    s1.render_all()
    a, _, _ = s1.download()
    s1.reset()
    s1.render_all()
    b, _, _ = s1.download()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_progressive_passes_equal_single_pass():
    sc = configs.cornell(width=40, height=40, spp=24).build()
    s = sc.session()
    s.render(0, 24)
    a, _, _ = s.download()
    s.reset()
    for off in range(0, 24, 5):
        s.render(off, min(5, 24 - off))
    b, _, _ = s.download()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_instanced_mini_parity(orc, asset_dir):
    w, h, spp = 96, 54, 32
    sc = configs.instanced(asset_dir, width=w, height=h, spp=spp, instances=96, subdiv=3).build()
    _, gpu, ref = _render_pair(sc, w, h, spp, orc)
    err = sd.rel_rmse(gpu, ref)
    assert err < 0.03, err


def test_volumes_mini_parity(orc, asset_dir):
    w, h, spp = 64, 36, 32
    sc = configs.volumes(asset_dir, width=w, height=h, spp=spp, env_w=256, env_h=128).build()
    _, gpu, ref = _render_pair(sc, w, h, spp, orc)
    err = sd.rel_rmse(gpu, ref)
    assert err < 0.05, err
