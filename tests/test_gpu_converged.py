"""Converged-image parity: at 4,096 spp the HIP film must match the CPU
oracle within 1% relative RMSE (over pixels x RGB of the linear film), on
each of the three scene families. Both renderers draw the same PCG32
streams, so the residual is where float differences (transcendentals,
contraction) flip a branch, not Monte Carlo noise. The oracle renders every
`stride`-th 16x16 tile to bound its time; only those tiles are compared."""

import numpy as np
import pytest

import sundog_amd as sd
from sundog_amd import configs

pytestmark = pytest.mark.gpu

SPP = 4096


def _tile_mask(w, h, stride):
    tx = (w + 15) // 16
    m = np.zeros((h, w), bool)
    for t in range(0, tx * ((h + 15) // 16), stride):
        bx, by = (t % tx) * 16, (t // tx) * 16
        m[by:by + 16, bx:bx + 16] = True
    return m


@pytest.mark.parametrize("name,w,h,stride", [
    ("cornell", 64, 64, 1),
    ("instanced", 64, 36, 1),
    ("volumes", 48, 32, 2),
])
def test_converged_parity_4096spp(name, w, h, stride, orc, asset_dir):
    if name == "cornell":
        scene = configs.cornell(width=w, height=h, spp=SPP)
    elif name == "instanced":
        scene = configs.instanced(asset_dir, width=w, height=h, spp=SPP, instances=64, subdiv=3)
    else:
        scene = configs.volumes(asset_dir, width=w, height=h, spp=SPP, env_w=256, env_h=128)
    sc = scene.build()
    sess = sc.session()
    sess.render_all(chunk=256)
    gpu, _, _ = sess.download()
    ref, _, _, _ = orc.render(sc.flatten(), w, h, 0, SPP, tile_stride=stride)
    m = _tile_mask(w, h, stride)
    err = sd.rel_rmse(gpu[m], ref[m])
    print("%s %dx%d @ %d spp: relRMSE %.5f" % (name, w, h, SPP, err))
    assert np.isfinite(gpu).all()
    assert err < 0.01, err


@pytest.mark.parametrize("name,stride", [("cornell", 4), ("instanced", 1021), ("volumes", 601)])
def test_converged_parity_4096spp_full_configs(name, stride, orc, asset_dir):
    """The benchmark configurations at full size: the GPU renders the whole
    frame at 4,096 spp (the Cornell box 256x256, the 4,096-instance frame
    1920x1080, the volumes frame 1280x720 with its 2048x1024 sky); the
    oracle renders every `stride`-th 16x16 tile (64, 8 and 6 tiles)."""
    if name == "cornell":
        scene = configs.cornell(spp=SPP)
        w, h = 256, 256
    elif name == "instanced":
        scene = configs.instanced(asset_dir, spp=SPP)
        w, h = 1920, 1080
    else:
        scene = configs.volumes(asset_dir, spp=SPP)
        w, h = 1280, 720
    sc = scene.build()
    sess = sc.session()
    sess.render_all(chunk=16)   # 16-spp passes: the session splits them into sample groups
    gpu, _, _ = sess.download()
    sess.close()
    print("%s %dx%d @ %d spp rendered on the GPU" % (name, w, h, SPP), flush=True)
    ref, _, _, _ = orc.render(sc.flatten(), w, h, 0, SPP, tile_stride=stride)
    m = _tile_mask(w, h, stride)
    err = sd.rel_rmse(gpu[m], ref[m])
    print("%s %dx%d @ %d spp: relRMSE %.5f over %d px" % (name, w, h, SPP, err, int(m.sum())), flush=True)
    assert np.isfinite(gpu).all()
    assert m.sum() >= 6 * 256
    assert err < 0.01, err
