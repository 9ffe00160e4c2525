"""Serving harness: resident scenes, Poisson request arrivals, continuous
batching on the GPU, latency percentiles, and the serial per-request
baseline on the same resident scenes.

Workload (defaults): four resident scenes (Cornell box; instanced
icospheres with 4,096 and 16,384 instances; volumes + environment), requests
arriving as a Poisson process at a fixed rate with a seeded generator, each
with a resolution drawn from {256^2, 512^2, 1024x576, 1920x1080}, spp
log-uniform in [16, 1024] and its own seed.

    python -m sundog_amd.serving --requests 64 --qps 4
"""

import argparse
import ctypes
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

from . import _lib
from . import configs

RESOLUTIONS = [(256, 256), (512, 512), (1024, 576), (1920, 1080)]


class Server(object):
    """ctypes wrapper of sundog_server_*."""

    def __init__(self, budget_bytes=64 << 30, pass_spp=16, tile=16, lib=None):
        self.lib = lib or _lib.load()
        self.ptr = self.lib.sundog_server_create(int(budget_bytes), pass_spp, tile)
        if not self.ptr:
            raise _lib.SundogError(_lib.last_error(self.lib))

    def close(self):
        if self.ptr:
            self.lib.sundog_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, handle, **overrides):
        opt = _lib.default_options(**overrides)
        sid = self.lib.sundog_server_add_scene(self.ptr, handle.ptr, ctypes.byref(opt))
        if sid < 0:
            raise _lib.SundogError(_lib.last_error(self.lib))
        return sid

    def tune(self, key, value):
        _lib.check(self.lib.sundog_server_set_tuning(self.ptr, key.encode(), int(value)), self.lib)

    def submit(self, scene, width, height, spp, seed=-1):
        rid = self.lib.sundog_server_submit(self.ptr, scene, width, height, spp, seed)
        if rid < 0:
            raise _lib.SundogError(_lib.last_error(self.lib))
        return rid

    def poll(self, max_n=256):
        arr = (_lib.RequestInfo * max_n)()
        n = self.lib.sundog_server_poll(self.ptr, arr, max_n)
        if n < 0:
            raise _lib.SundogError(_lib.last_error(self.lib))
        return [{k: getattr(arr[i], k) for k, _ in _lib.RequestInfo._fields_ if k != "pad"} for i in range(n)]

    def fetch(self, rid, width, height):
        b = np.zeros(width * height * 4, np.float32)
        _lib.check(self.lib.sundog_server_fetch(self.ptr, rid, b.ctypes.data_as(_lib.P_f), None, None), self.lib)
        return b.reshape(height, width, 4)

    def release(self, rid):
        _lib.check(self.lib.sundog_server_release(self.ptr, rid), self.lib)

    def wait_idle(self, timeout=0.0):
        _lib.check(self.lib.sundog_server_wait_idle(self.ptr, timeout), self.lib)

    def stats(self):
        st = _lib.ServerStats()
        _lib.check(self.lib.sundog_server_get_stats(self.ptr, ctypes.byref(st)), self.lib)
        return {k: getattr(st, k) for k, _ in st._fields_}

    def autotune(self, scene, width=512, height=512, spp=32):
        """Picks waves, stack, tile, queue depth and graph passes by measuring
        a calibration request; call before the first submit."""
        r = _lib.TuneResult()
        _lib.check(self.lib.sundog_server_autotune(self.ptr, scene, width, height, spp, ctypes.byref(r)),
                   self.lib)
        return {k: getattr(r, k) for k, _ in r._fields_}

    def render_serial(self, scene, width, height, spp, seed=-1, want_image=False):
        ms = ctypes.c_double(0)
        b = np.zeros(width * height * 4, np.float32) if want_image else None
        _lib.check(self.lib.sundog_server_render_serial(
            self.ptr, scene, width, height, spp, seed,
            b.ctypes.data_as(_lib.P_f) if b is not None else None, ctypes.byref(ms)), self.lib)
        return ms.value, (b.reshape(height, width, 4) if b is not None else None)


def make_requests(n, qps, seed=1, scenes=4, spp_lo=16, spp_hi=1024, resolutions=RESOLUTIONS):
    """Seeded Poisson arrivals; resolution uniform over `resolutions`, spp
    = round(exp(U[ln spp_lo, ln spp_hi])), one seed per request."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / qps)
        w, h = resolutions[rng.integers(len(resolutions))]
        spp = int(round(math.exp(rng.uniform(math.log(spp_lo), math.log(spp_hi)))))
        reqs.append({"arrival_s": t, "scene": int(rng.integers(scenes)), "width": w, "height": h,
                     "spp": spp, "seed": int(rng.integers(1 << 31))})
    return reqs


def build_scenes(asset_dir, small=False):
    """The four resident scenes of the serving workload."""
    if small:
        return [configs.cornell(),
                configs.instanced(asset_dir, instances=64, subdiv=3),
                configs.instanced(asset_dir, instances=256, subdiv=3),
                configs.volumes(asset_dir, env_w=256, env_h=128)]
    return [configs.cornell(),
            configs.instanced(asset_dir),
            configs.instanced(asset_dir, instances=16384),
            configs.volumes(asset_dir)]


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * p / 100.0
    f, c = int(math.floor(k)), int(math.ceil(k))
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def run(requests, server, scene_ids, realtime=True):
    """Submits requests at their arrival times, waits for all; returns
    per-request records and the makespan."""
    t0 = time.perf_counter()
    ids = {}
    for r in requests:
        if realtime:
            delay = r["arrival_s"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        ids[server.submit(scene_ids[r["scene"]], r["width"], r["height"], r["spp"], r["seed"])] = r
    server.wait_idle()
    makespan = time.perf_counter() - t0
    done = server.poll(len(requests) + 16)
    return done, makespan


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--qps", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--waves", type=int, default=4)
    ap.add_argument("--stack", type=int, default=16)
    ap.add_argument("--graph-passes", type=int, default=4)
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shape with the autotuner on a calibration request of scene 1")
    ap.add_argument("--autotune-size", default="1920x1080x32",
                    help="calibration request WIDTHxHEIGHTxSPP (large enough that tails do not dominate)")
    ap.add_argument("--serial-waves", type=int, default=0,
                    help="occupancy variant of the per-request megakernel baseline (0: picked per "
                         "request by timing its first passes, as a session does)")
    ap.add_argument("--pass-spp", type=int, default=16)
    ap.add_argument("--groups", type=int, default=0,
                    help="sample groups per pass: 0 splits small batches to fill the resident waves, 1 never")
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--spp-max", type=int, default=1024)
    ap.add_argument("--small-scenes", action="store_true")
    ap.add_argument("--no-serial", action="store_true")
    ap.add_argument("--asset-dir", default=os.path.join(tempfile.gettempdir(), "sundog_assets"))
    args = ap.parse_args(argv)

    scenes = build_scenes(args.asset_dir, small=args.small_scenes)
    handles = [s.build() for s in scenes]
    srv = Server(budget_bytes=int(args.budget_gb * (1 << 30)), pass_spp=args.pass_spp, tile=args.tile)
    sids = [srv.add_scene(h) for h in handles]
    srv.tune("waves", args.waves)
    srv.tune("stack", args.stack)
    srv.tune("graph_passes", args.graph_passes)
    srv.tune("serial_waves", args.serial_waves)
    srv.tune("groups", args.groups)
    cw, ch, cs = (int(v) for v in args.autotune_size.split("x"))
    tuned = srv.autotune(sids[min(1, len(sids) - 1)], cw, ch, cs) if args.autotune else None
    reqs = make_requests(args.requests, args.qps, seed=args.seed, spp_hi=args.spp_max)
    total_paths = sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    print("serving: %d requests, %.3g paths, qps %.2f" % (len(reqs), total_paths, args.qps), flush=True)
    done, makespan = run(reqs, srv, sids)
    lat = [d["t_done_ms"] - d["t_submit_ms"] for d in done if d["status"] == 2]
    st = srv.stats()
    out = {
        "requests": len(reqs), "completed": len(lat),
        "paths": total_paths,
        "makespan_s": makespan,
        "served_paths_per_s": total_paths / makespan,
        "served_kernel_paths_per_s": st["paths"] / (st["kernel_ms"] * 1e-3) if st["kernel_ms"] else None,
        "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
        "batches": st["batches"], "graph_launches": st["graph_launches"], "tile": st["tile_size"],
        "waves": st["waves"], "stack": st["stack"], "blocks_per_cu": st["blocks_per_cu"],
        "graph_passes": st["graph_passes"], "pass_spp": st["pass_spp"], "autotune": tuned,
        "mrays_per_s": st["rays_traced"] / makespan / 1e6,
    }
    if not args.no_serial:
        ser_ms = 0.0
        for r in reqs:
            ms, _ = srv.render_serial(sids[r["scene"]], r["width"], r["height"], r["spp"], r["seed"])
            ser_ms += ms
        out["serial_paths_per_s"] = total_paths / (ser_ms * 1e-3)
        out["serial_waves"] = args.serial_waves
        out["speedup_vs_serial"] = out["served_paths_per_s"] / out["serial_paths_per_s"]
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
