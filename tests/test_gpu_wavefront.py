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
    """The megakernel's occupancy variants (2 to 5 waves per SIMD, different
    register budgets and spills) give the same bits."""
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


@pytest.mark.parametrize("name", ["cornell", "instanced_mini", "volumes_mini"])
def test_sample_groups_bit_identical(name, asset_dir):
    """Sample groups split a pass's samples over several blocks per tile and
    fold them in a separate kernel; every group count, even splits and
    uneven ones (3 or 7 groups of a 12-sample pass, 16 groups of 12 samples
    leaving some groups empty), gives the same bits as one group."""
    sc = _scene(name, asset_dir)
    spp = sc.session().spp
    ref = None
    for groups in (1, 2, 3, 7, 16):
        s = sc.session()
        s.tune("waves", 3)
        s.tune("groups", groups)
        for off in range(0, spp, 12):
            s.render(off, min(12, spp - off))
        films = s.download()
        assert s.stats()["sample_groups"] == groups
        if ref is None:
            ref = films
        else:
            for a, b in zip(ref, films):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), groups


@pytest.mark.parametrize("name", ["cornell", "instanced_mini"])
def test_autotune_picks_a_variant_and_keeps_the_bits(name, asset_dir):
    """sundog_session_autotune times the variants on real passes, keeps one
    and clears the film; renders afterwards (and renders in auto mode, whose
    first passes are the timed trials) equal a fixed-variant render."""
    sc = _scene(name, asset_dir)
    spp = sc.session().spp
    fixed = sc.session()
    fixed.tune("waves", 2)
    fixed.render(0, spp)
    ref = fixed.download()
    s = sc.session()
    w = s.autotune(4)
    assert w in (2, 3, 4, 5)
    st = s.stats()
    assert st["waves"] == w and st["paths"] == 0
    assert st["sample_groups"] in (1, 2, 3, 4)    # candidates up to the pass's 4 samples
    assert not any(f.any() for f in s.download())      # film cleared
    s.render(0, spp)
    for a, b in zip(ref, s.download()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    auto = sc.session()                                   # auto mode, no explicit call
    for off in range(0, spp, 2):
        auto.render(off, min(2, spp - off))
    assert auto.stats()["waves"] in (0, 2, 3, 4, 5)
    for a, b in zip(ref, auto.download()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
