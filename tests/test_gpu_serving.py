"""GPU serving: continuous batching through the persistent tile-queue
kernel. Each request's film must be bit-identical rendered alone (session
megakernel) and in a mix of concurrent requests; admission must respect the
HBM budget; oversize requests are rejected."""

import numpy as np
import pytest

from sundog_amd import configs
from sundog_amd.serving import Server

pytestmark = pytest.mark.gpu


def _scenes(asset_dir):
    return [configs.cornell(width=64, height=64, spp=32).build(),
            configs.instanced(asset_dir, width=64, height=36, spp=16, instances=32, subdiv=2).build()]


@pytest.mark.parametrize("tile,graph_passes,stack", [(8, 4, 16), (16, 4, 8), (32, 4, 24), (16, 1, 16)])
def test_mix_equals_alone(asset_dir, tile, graph_passes, stack):
    """Continuous batching, with passes captured into graphs (graph_passes >
    1) or launched one by one, and with spill-heavy (8) or roomy (24) LDS
    stacks: every film equals the request rendered alone."""
    hs = _scenes(asset_dir)
    srv = Server(budget_bytes=1 << 30, pass_spp=8, tile=tile)
    srv.tune("graph_passes", graph_passes)
    srv.tune("stack", stack)
    sids = [srv.add_scene(h) for h in hs]
    reqs = [(sids[0], 48, 40, 24, 5), (sids[1], 64, 36, 16, 9), (sids[0], 33, 17, 40, 2),
            (sids[1], 20, 70, 8, 1), (sids[0], 64, 64, 32, 1)]
    ids = [srv.submit(*r) for r in reqs]
    srv.wait_idle(120.0)
    done = {d["id"]: d for d in srv.poll()}
    assert len(done) == len(reqs) and all(d["status"] == 2 for d in done.values())
    for rid, (sc, w, h, spp, seed) in zip(ids, reqs):
        mix = srv.fetch(rid, w, h)
        _, alone = srv.render_serial(sc, w, h, spp, seed, want_image=True)
        assert np.array_equal(mix.view(np.uint32), alone.view(np.uint32)), (rid, w, h, spp)
        assert done[rid]["t_done_ms"] >= done[rid]["t_admit_ms"] >= done[rid]["t_submit_ms"]
    st = srv.stats()
    assert st["paths"] == sum(w * h * spp for _, w, h, spp, _ in reqs)


def test_served_film_matches_session_render(asset_dir):
    h = configs.cornell(width=40, height=32, spp=24).build()
    sess = h.session()
    sess.render_all()
    ref, _, _ = sess.download()
    srv = Server(budget_bytes=1 << 28, pass_spp=5, tile=8)
    sid = srv.add_scene(h)
    rid = srv.submit(sid, 40, 32, 24, 1)
    srv.wait_idle(60.0)
    out = srv.fetch(rid, 40, 32)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_admission_by_footprint_and_rejection(asset_dir):
    h = configs.cornell(width=32, height=32, spp=4).build()
    budget = 64 * 64 * 48 + 1000          # one 64x64 film fits, not two
    srv = Server(budget_bytes=budget, pass_spp=4, tile=16)
    sid = srv.add_scene(h)
    a = srv.submit(sid, 64, 64, 16, 1)
    b = srv.submit(sid, 64, 64, 16, 2)
    c = srv.submit(sid, 128, 128, 4, 3)     # larger than the whole budget
    srv.wait_idle(60.0)
    rec = {d["id"]: d for d in srv.poll()}
    assert rec[c]["status"] == 3
    assert rec[a]["status"] == 2 and rec[b]["status"] == 2
    # b could only be admitted after a released its film
    assert rec[b]["t_admit_ms"] >= rec[a]["t_done_ms"]
    st = srv.stats()
    assert st["used_bytes"] == 0


def test_autotune_then_serve(asset_dir):
    hs = _scenes(asset_dir)
    srv = Server(budget_bytes=1 << 30, pass_spp=8, tile=16)
    sids = [srv.add_scene(h) for h in hs]
    res = srv.autotune(sids[1], 64, 36, 16)
    assert res["tile"] in (8, 16, 32) and res["waves"] in (2, 3, 4, 5) and 4 <= res["stack"] <= 24
    assert res["trials"] >= 10 and res["best_paths_per_s"] >= res["start_paths_per_s"] > 0
    st = srv.stats()
    assert st["paths"] == 0 and st["tile_size"] == res["tile"] and st["stack"] == res["stack"]
    rid = srv.submit(sids[0], 40, 24, 20, 3)
    srv.wait_idle(60.0)
    out = srv.fetch(rid, 40, 24)
    _, alone = srv.render_serial(sids[0], 40, 24, 20, 3, want_image=True)
    assert np.array_equal(out.view(np.uint32), alone.view(np.uint32))


@pytest.mark.parametrize("groups,tile", [(1, 8), (3, 8), (8, 16), (0, 32)])
def test_sample_groups_keep_films(asset_dir, groups, tile):
    """A pass split into sample groups (several work items per tile, each
    rendering a share of the pass's samples into per-pool regions that a
    fold kernel folds in sample order), never split (1), or split
    automatically (0: small batches fill the resident waves): every film
    equals the request rendered alone."""
    hs = _scenes(asset_dir)
    srv = Server(budget_bytes=1 << 30, pass_spp=8, tile=tile)
    srv.tune("groups", groups)
    sids = [srv.add_scene(h) for h in hs]
    reqs = [(sids[0], 48, 40, 24, 5), (sids[1], 64, 36, 13, 9), (sids[0], 33, 17, 40, 2)]
    ids = [srv.submit(*r) for r in reqs]
    srv.wait_idle(120.0)
    done = {d["id"]: d for d in srv.poll()}
    assert len(done) == len(reqs) and all(d["status"] == 2 for d in done.values())
    for rid, (sc, w, h, spp, seed) in zip(ids, reqs):
        mix = srv.fetch(rid, w, h)
        _, alone = srv.render_serial(sc, w, h, spp, seed, want_image=True)
        assert np.array_equal(mix.view(np.uint32), alone.view(np.uint32)), (groups, rid, w, h, spp)
    assert srv.stats()["paths"] == sum(w * h * spp for _, w, h, spp, _ in reqs)
