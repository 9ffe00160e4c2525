"""GPU BVH builds (the LBVH over both levels, the LBVH mesh levels under a
host-SAH top level over GPU instance boxes, and PLOC clustering over both
levels or under a host-SAH top level): closest hits through the rebuilt two-level
BVH equal brute force bit-exactly (t bits, instance, primitive), and films
equal the host-SAH BVH's films bit-exactly where shadow rays are opaque
(closest hits and opaque occlusion do not depend on the tree)."""

import numpy as np
import pytest

from sundog_amd import bvhinfo, configs

pytestmark = pytest.mark.gpu


def _rays(n, seed, eye_lo, eye_hi, tgt_lo, tgt_hi):
    rng = np.random.default_rng(seed)
    o = rng.uniform(eye_lo, eye_hi, (n, 3))
    d = rng.uniform(tgt_lo, tgt_hi, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = 1e16
    return r


def _scene(name, asset_dir, **kw):
    if name == "cornell":
        return configs.cornell(**kw)
    if name == "instanced_mini":
        return configs.instanced(asset_dir, instances=96, subdiv=3, **kw)
    return configs.instanced(asset_dir, instances=2048, subdiv=4, **kw)


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
@pytest.mark.parametrize("name", ["cornell", "instanced_mini", "instanced_2k"])
def test_lbvh_closest_hit_equals_brute_force(name, mode, asset_dir):
    sc = _scene(name, asset_dir, width=32, height=32, spp=1).build()
    sess = sc.session()
    sess.tune("bvh", mode)
    st = sess.stats()
    assert st["bvh_builder"] == mode and st["bvh_gpu_ms"] > 0 and st["tlas_nodes"] >= 1
    if name == "cornell":
        rays = _rays(4096, 3, -0.9, 0.9, -0.9, 0.9)
    else:
        rays = _rays(4096, 4, [-40, 2, -40], [40, 30, 60], [-50, 0, -50], [50, 3, 50])
    t, i, p = sess.trace(rays)
    tb, ib, pb = sess.trace(rays, brute=True)
    assert int((i >= 0).sum()) > len(rays) // 4
    assert np.array_equal(i, ib) and np.array_equal(p, pb)
    assert np.array_equal(t.view(np.uint32), tb.view(np.uint32))


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
@pytest.mark.parametrize("name", ["cornell", "instanced_mini"])
def test_lbvh_film_equals_sah_film(name, mode, asset_dir):
    sc = _scene(name, asset_dir, width=48, height=32, spp=8).build()
    a = sc.session()
    a.render_all()
    fa, _, _ = a.download()
    b = sc.session()
    b.tune("bvh", mode)
    b.render_all()
    fb, _, _ = b.download()
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_lbvh_full_instanced_scene_equals_sah(mode, asset_dir):
    """The full 4,096-instance scene (20,480-triangle mesh): the GPU sort and
    scan run over several blocks (the mesh's keys span five sort blocks and
    its 20,479 internal nodes ten scan blocks); closest hits through the
    rebuilt BVH equal those through the host-SAH BVH bit-exactly."""
    sc = configs.instanced(asset_dir, width=32, height=32, spp=1).build()
    a = sc.session()
    b = sc.session()
    b.tune("bvh", mode)
    assert b.stats()["bvh_builder"] == mode
    rays = _rays(8192, 5, [-60, 1, -60], [60, 40, 70], [-50, 0, -50], [50, 4, 50])
    ta, ia, pa = a.trace(rays)
    tb, ib, pb = b.trace(rays)
    assert int((ia >= 0).sum()) > len(rays) // 2
    assert np.array_equal(ia, ib) and np.array_equal(pa, pb)
    assert np.array_equal(ta.view(np.uint32), tb.view(np.uint32))


@pytest.mark.parametrize("mode", [0, 1, 3, 4])
def test_builds_are_deterministic_and_costed(mode, asset_dir):
    """Two builds of the same scene (the host SAH build, or a GPU rebuild)
    give the same node and triangle bytes, and the tree's SAH cost is
    reported from those bytes."""
    sc = configs.instanced(asset_dir, instances=2048, subdiv=4, width=32, height=32, spp=1).build()
    recs = []
    for _ in range(2):
        s = sc.session()
        if mode:
            s.tune("bvh", mode)
        assert s.stats()["bvh_builder"] == mode
        recs.append(s.bvh())
    a, b = recs
    assert a["num_nodes"] == b["num_nodes"] and a["nodes"] == b["nodes"]
    assert np.array_equal(a["tris"].view(np.uint32), b["tris"].view(np.uint32))
    # every triangle of the mesh once, in leaf order
    ids = np.sort(a["tris"][:, 3].view(np.int32))
    assert np.array_equal(ids, np.arange(a["num_tris"]))
    summ = bvhinfo.summary(a)
    print("bvh mode %d: %s" % (mode, summ))
    assert summ["tlas_sah"] > 1.0 and len(summ["blas_sah"]) == 1
    assert all(v > 1.0 for v in summ["blas_sah"].values())


@pytest.mark.parametrize("mode", [0, 1, 3, 4])
def test_coincident_instances_closest_hit_equals_brute_force(mode, asset_dir):
    """Five instances of one mesh at the same transform and two sphere
    quadrics at the same place (identical boxes: every clustering distance
    ties, and every hit on them ties in t across instances) plus two
    instances elsewhere: closest hits through each build equal brute force
    bit-exactly, the lowest instance winning each tie."""
    import os
    from sundog_amd.configs import icosphere, rotate_y, scale, translate, write_obj
    from sundog_amd.scenelib import Scene
    obj = os.path.join(asset_dir, "icosphere_s2_tie.obj")
    if not os.path.exists(obj):
        v, f = icosphere(2, amplitude=0.1, seed=7)
        tmp = obj + ".tmp%d" % os.getpid()
        write_obj(tmp, v, f)
        os.replace(tmp, obj)
    s = Scene()
    s.render(width=16, height=16, spp=1, max_depth=4, seed=1)
    s.camera(lookfrom=[0, 2, 8], lookat=[0, 0, 0], vfov=40)
    s.mesh("ico", obj, normals="smooth")
    s.lambert("m", color=[0.5, 0.5, 0.5])
    for _ in range(5):
        s.add("mesh:ico", "m", scale(1.0, 0.8, 1.2), rotate_y(30.0), translate(0.0, 0.5, 0.0))
    s.add("sphere", "m", scale(0.7), translate(2.5, 0.5, 0.0))
    s.add("sphere", "m", scale(0.7), translate(2.5, 0.5, 0.0))
    s.add("mesh:ico", "m", translate(-2.5, 0.5, 0.0))
    s.add("mesh:ico", "m", scale(0.5), translate(0.0, 0.5, -3.0))
    sess = s.build().session()
    if mode:
        sess.tune("bvh", mode)
    assert sess.stats()["bvh_builder"] == mode
    rays = _rays(4096, 9, [-4, -1, 5], [4, 3, 9], [-3.5, -0.5, -3.5], [3.5, 1.5, 1.5])
    t, i, p = sess.trace(rays)
    tb, ib, pb = sess.trace(rays, brute=True)
    assert int((i >= 0).sum()) > len(rays) // 4
    assert set(np.unique(i[i >= 0])) <= {0, 5, 7, 8}   # ties go to the first of the coincident instances
    assert np.array_equal(i, ib) and np.array_equal(p, pb)
    assert np.array_equal(t.view(np.uint32), tb.view(np.uint32))


def test_bvh_export_checks_buffer_sizes(asset_dir):
    """sundog_session_bvh fills the layout with NULL buffers and refuses a
    node or triangle buffer smaller than the layout asks for."""
    import ctypes
    from sundog_amd import _lib
    sc = configs.instanced(asset_dir, instances=8, subdiv=2, width=16, height=16, spp=1).build()
    sess = sc.session()
    lay = _lib.BvhLayout()
    lib = sess.lib
    assert lib.sundog_session_bvh(sess.ptr, ctypes.byref(lay), None, 0, None, 0, None, 0) == 0
    assert lay.num_nodes >= 2 and lay.tlas_nodes >= 1 and lay.num_tris == 320 and lay.num_instances == 9
    small = np.zeros(lay.num_nodes * 32 - 1, np.float32)
    assert lib.sundog_session_bvh(sess.ptr, ctypes.byref(lay), small.ctypes.data_as(ctypes.c_void_p),
                                  small.nbytes, None, 0, None, 0) != 0
    assert "node buffer" in _lib.last_error(lib)
    tris = np.zeros((lay.num_tris - 1, 12), np.float32)
    assert lib.sundog_session_bvh(sess.ptr, ctypes.byref(lay), None, 0,
                                  tris.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), tris.nbytes, None, 0) != 0
    assert "triangle buffer" in _lib.last_error(lib)
