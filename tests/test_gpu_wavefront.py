"""The wavefront variant (path state in HBM, queued closest-hit and shadow
queries, material-sorted shading) against the megakernel: both run the same
path stages, so every film must be bit-identical, and so must the ray
counts."""

import numpy as np
import pytest

from sundog_amd import configs

pytestmark = pytest.mark.gpu


def _scene(name, asset_dir):
    if name == "cornell":
        return configs.cornell(width=48, height=40, spp=24).build()
    if name == "instanced_mini":
        return configs.instanced(asset_dir, width=96, height=54, spp=12, instances=96, subdiv=3).build()
    return configs.volumes(asset_dir, width=64, height=36, spp=12, env_w=256, env_h=128).build()


def _render(sc, kernel, passes):
    s = sc.session()
    s.set_kernel(kernel)
    s.count_rays(True)
    for off, cnt in passes:
        s.render(off, cnt)
    films = s.download()
    return films, s.stats()["rays_traced"]


@pytest.mark.parametrize("name", ["cornell", "instanced_mini", "volumes_mini"])
def test_wavefront_equals_megakernel(name, asset_dir):
    sc = _scene(name, asset_dir)
    spp = sc.session().spp
    passes = [(0, spp)]
    mk, rays_mk = _render(sc, "megakernel", passes)
    wf, rays_wf = _render(sc, "wavefront", passes)
    for a, b in zip(mk, wf):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert rays_mk == rays_wf


def test_wavefront_progressive_passes(asset_dir):
    sc = _scene("cornell", asset_dir)
    (one, _, _), _ = _render(sc, "wavefront", [(0, 24)])
    (many, _, _), _ = _render(sc, "wavefront", [(0, 5), (5, 7), (12, 1), (13, 11)])
    assert np.array_equal(one.view(np.uint32), many.view(np.uint32))


@pytest.mark.parametrize("name", ["cornell", "instanced_mini", "volumes_mini"])
def test_occupancy_variants_bit_identical(name, asset_dir):
    """The megakernel's occupancy variants (film running means in registers
    at 2 waves per SIMD, folded in memory at 3 to 5) give the same bits."""
    sc = _scene(name, asset_dir)
    spp = sc.session().spp
    ref = None
    for waves in (2, 3, 4, 5):
        s = sc.session()
        s.tune("waves", waves)
        s.render(0, spp)
        films = s.download()
        if ref is None:
            ref = films
        else:
            for a, b in zip(ref, films):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), waves
