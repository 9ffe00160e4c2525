"""The allocation-free closest-hit query (sundog_session_trace_noalloc)
against the allocating one (sundog_session_trace): bit-identical (t bits,
instance, primitive) on two calls in a row, the second with fewer rays, on
rays and outputs that live in device memory (torch tensors as plumbing)."""

import numpy as np
import pytest

import sundog_amd as sd
from sundog_amd import configs

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform([-60, 2, -60], [60, 30, 60], (n, 3))
    t = rng.uniform([-50, 0, -50], [50, 3, 50], (n, 3))
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = 1e16
    return r


def test_trace_noalloc_matches_trace(asset_dir):
    import torch
    sc = configs.instanced(asset_dir, width=32, height=18, spp=1, instances=256, subdiv=3).build()
    sess = sc.session()
    lib = sess.lib
    for n, seed in ((40000, 1), (5000, 2)):
        print("trace_noalloc: %d rays" % n, flush=True)
        rays = _rays(n, seed)
        t_ref, i_ref, p_ref = sess.trace(rays)
        assert (i_ref >= 0).mean() > 0.2
        d_rays = torch.from_numpy(rays).to("cuda")
        d_t = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        d_i = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        d_p = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = lib.sundog_session_trace_noalloc(sess.ptr, d_rays.data_ptr(), n, d_t.data_ptr(), d_i.data_ptr(),
                                              d_p.data_ptr())
        assert rc == 0, sd.last_error() if hasattr(sd, "last_error") else rc
        sess.sync()
        t = d_t.cpu().numpy()
        i = d_i.cpu().numpy()
        p = d_p.cpu().numpy()
        assert np.array_equal(i, i_ref)
        assert np.array_equal(p, p_ref)
        assert np.array_equal(t.view(np.uint32), t_ref.view(np.uint32))
