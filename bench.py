"""Benchmark: instanced-icosphere path tracing on one MI355X.

Workload (the headline configuration): 4,096 instances of a displaced
20,480-triangle icosphere on a 100 x 100 plane, 50% GGX metal / 50%
lambert, 3 soft point lights + 1 distant light, 1920x1080, 256 spp, max
depth 12, fp32 film. Scene data (two-level BVH, instances, materials,
film) is resident in HBM before timing starts.

A step is one progressive pass of 16 samples per pixel over the full frame
(the reference's launch granularity); 16 steps are the full 256-spp frame.
value = paths/s over the timed steps (wall clock, synchronized on both
sides). The roofline entry prices the render kernel's algorithmic traversal
FLOPs (12 per box test, 40 per triangle test, 30 per quadric test, 30 per
instance transform; counted by an instrumented pass) against the FP32
vector peak; the kernel time comes from HIP events on the session stream.
cpu_baseline times the CPU oracle on a bounded sample of the same frame,
on the host cores given to this job. A short run of the serving workload
(four resident scenes, Poisson arrivals, continuous batching in the
persistent kernel) adds the served and serial-megakernel rates, their
ratio and the p50 / p99 request latencies under "serving". The two
reference gallery scenes that need no asset files (01-marble-run and
02-cornell-lume, replayed from their recorded call programs) are rendered
as the reference's own benchmark renders them (1920x1080, 64 spp, render
loop wall time and Mrays/s from sundog_render's stats file) and reported
under "reference_scenes" beside the figures the reference publishes for them.
"""

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "paths/s & Mrays/s at 1080p; p50/p99 request latency at ≤1% RMSE"
# MI355X published FP32 vector peak (157.3 TFLOP/s counts packed-FP32 FMAs)
# and the non-packed issue peak (256 CUs x 64 lanes x 2 flops per FMA x
# 2.4 GHz); this build emits no packed FP32, so the second bounds it.
FP32_PEAK_TFLOPS = 157.3
FP32_PEAK_NONPACKED_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE_NAME = "profiles/k_render_instanced_1080p.json"
TRAFFIC_PROFILE = os.path.join(ROOT, TRAFFIC_PROFILE_NAME)


def counter_profile(kernel):
    """The committed counter profile of this workload's kernel (FETCH_SIZE /
    WRITE_SIZE traffic, VALU busy and lane utilisation), or {}."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return {}
    return prof if prof.get("kernel") == kernel else {}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE counter profile of this workload, or None."""
    return counter_profile(kernel).get("traffic_bytes")


def copy_probe(mib):
    """A plain copy kernel of known size: reads and writes mib MiB per launch."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 4}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=16)
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--workload", default="instanced", choices=["instanced", "cornell", "volumes"],
                    help="the headline instanced-icosphere frame (default), or the Cornell-box and "
                         "volumes-and-environment frames")
    ap.add_argument("--waves", type=int, default=0, choices=[0, 2, 3, 4, 5],
                    help="minimum waves per SIMD the megakernel's register budget allows; 0 (default) "
                         "lets the session time the variants on real passes before the timed region")
    ap.add_argument("--stack", type=int, default=12,
                    help="traversal stack entries per lane kept in LDS (deeper entries spill to HBM)")
    ap.add_argument("--groups", type=int, default=0,
                    help="megakernel sample groups per tile (0: the session's choice by frame size)")
    ap.add_argument("--kernel", default="megakernel", choices=["megakernel", "wavefront"])
    ap.add_argument("--wf-waves", type=int, default=4, choices=[4, 5],
                    help="minimum waves per SIMD of the wavefront query kernels")
    ap.add_argument("--bvh", default="sah", choices=["sah", "lbvh", "lbvh-sah-top", "ploc", "ploc-sah-top"],
                    help="host binned-SAH build (default), the GPU LBVH rebuild of both levels, the GPU "
                         "LBVH for the meshes with a host SAH top level over GPU instance boxes, the GPU "
                         "PLOC rebuild of both levels, or GPU PLOC meshes under a host SAH top level")
    ap.add_argument("--ploc-radius", type=int, default=16,
                    help="neighbours searched on either side by the PLOC build (1..32)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0: the job's CPU share, OMP_NUM_THREADS capped by the affinity mask)")
    ap.add_argument("--serving-requests", type=int, default=24,
                    help="requests of the serving run reported under 'serving' (0: skip it)")
    ap.add_argument("--serving-qps", type=float, default=4.0,
                    help="arrival rate of the saturating serving run (throughput and served/serial ratio)")
    ap.add_argument("--serving-qps-sustained", type=float, default=1.0,
                    help="arrival rate of a second run of the same requests below the served capacity, "
                         "whose p50 / p99 measure service rather than queue growth (0: skip it)")
    ap.add_argument("--no-reference-scenes", action="store_true",
                    help="skip the reference gallery scenes reported under 'reference_scenes'")
    ap.add_argument("--cpu-tile-stride", type=int, default=8)
    ap.add_argument("--cpu-spp", type=int, default=128)
    ap.add_argument("--asset-dir", default=os.path.join(tempfile.gettempdir(), "sundog_assets"))
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB, 4 times "
                         "(the calibration kernel for FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(the render path would shard by pixel tiles, one set of tiles per GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import sundog_amd as sd
    from sundog_amd import configs

    t0 = time.time()
    if args.workload == "instanced":
        W, H, SPP, DEPTH = 1920, 1080, 256, 12
        scene = configs.instanced(args.asset_dir, width=W, height=H, spp=SPP, max_depth=DEPTH,
                                  instances=args.instances).build()
        workload = ("instanced icospheres: %d x 20,480-tri instances, %dx%d, %d spp, depth %d"
                    % (args.instances, W, H, SPP, DEPTH))
        data = "synthetic (seed-generated scene: icosphere OBJ + PCG32 layout)"
    elif args.workload == "cornell":
        W, H, SPP, DEPTH = 256, 256, 64, 8
        scene = configs.cornell(width=W, height=H, spp=SPP, max_depth=DEPTH).build()
        workload = "Cornell box (5 rects, lambert + glass spheres, rect light): %dx%d, %d spp, depth %d" % (
            W, H, SPP, DEPTH)
        data = "synthetic (the fixed Cornell layout, no assets)"
    else:
        W, H, SPP, DEPTH = 1280, 720, 512, 16
        scene = configs.volumes(args.asset_dir, width=W, height=H, spp=SPP, max_depth=DEPTH).build()
        workload = ("volumes + environment (2048x1024 sky, 5-octave flame box, water, rough glass, transparent "
                    "shadows): %dx%d, %d spp, depth %d" % (W, H, SPP, DEPTH))
        data = "synthetic (procedural 2048x1024 sky image written as .hdr, fixed layout)"
    sess = scene.session()
    sess.tune("waves", args.waves)
    sess.tune("stack", args.stack)
    sess.tune("groups", args.groups)
    sess.set_kernel(args.kernel)
    sess.tune("wf_waves", args.wf_waves)
    if args.bvh == "lbvh":
        sess.tune("bvh", 1)
    elif args.bvh == "lbvh-sah-top":
        sess.tune("bvh", 2)
    elif args.bvh in ("ploc", "ploc-sah-top"):
        sess.tune("ploc_radius", args.ploc_radius)
        sess.tune("bvh", 3 if args.bvh == "ploc" else 4)
    setup_s = time.time() - t0
    k = args.spp_per_step
    from sundog_amd import bvhinfo
    summ = bvhinfo.summary(sess.bvh())
    bvh_sah = {"top": summ["tlas_sah"], "meshes": list(summ["blas_sah"].values())}

    st0 = sess.stats()
    # a second session over the same scene times the build with the HIP
    # runtime already up (the first one's build_ms includes its start-up)
    warm = scene.session()
    warm_build_ms = warm.stats()["build_ms"]
    warm.close()
    # instrumented pass: traversal work per ray query (the megakernel's
    # counting variant; the wavefront variant runs the same traversal code)
    sess.set_kernel("megakernel")
    sess.count_rays(True)
    sess.count_work(True)
    sess.render(0, k)
    sess.sync()
    st = sess.stats()
    sess.set_kernel(args.kernel)
    queries = max(st["rays_traced"], 1)
    work = {"box_tests": 4 * st["node_visits"], "tri_tests": st["tri_tests"],
            "quad_tests": st["quad_tests"], "inst_transforms": st["inst_transforms"]}
    flops_instr = (12 * work["box_tests"] + 40 * work["tri_tests"] + 30 * work["quad_tests"]
                   + 30 * work["inst_transforms"])
    flops_per_query = flops_instr / queries
    sess.count_work(False)

    # occupancy variant: fixed, or timed on real passes before the warm-up
    # (sundog_session_autotune; the film is cleared afterwards)
    waves = args.waves
    if args.kernel == "megakernel" and (waves == 0 or args.groups == 0):
        waves = sess.autotune(k)   # times what is not fixed: the occupancy variant, the sample groups
    groups_used = sess.stats()["sample_groups"] if args.kernel == "megakernel" else 1

    # warmup, then the timed passes
    for i in range(args.warmup):
        sess.render((i * k) % SPP, k)
    sess.sync()
    sess.reset()
    sess.count_rays(True)
    sess.sync()
    t1 = time.perf_counter()
    for i in range(args.steps):
        sess.render((i * k) % SPP, k)
    sess.sync()
    wall = time.perf_counter() - t1
    st = sess.stats()
    paths = st["paths"]
    rays = st["rays_traced"]
    kernel_ms_avg = st["render_ms"] / args.steps
    flops_per_launch = flops_per_query * rays / args.steps
    achieved_tflops = flops_per_launch / (kernel_ms_avg * 1e-3) / 1e12
    kname = ("sdk::k_render<false, %d>" % waves if args.kernel == "megakernel"
             else "wavefront pass (k_wf_shade + k_wf_closest + k_wf_shadow)")
    beauty, _, _ = sess.download()
    finite = bool(np.isfinite(beauty).all())
    traffic = measured_traffic(kname) if args.workload == "instanced" else None

    out = {
        "metric": METRIC,
        "value": paths / wall,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": data,
        "config": {
            "workload": workload,
            "step": "one %d-spp progressive pass over the full frame" % k,
            "kernel": args.kernel,
            "waves_per_simd": waves,
            "waves_chosen_by": "autotune" if args.waves == 0 and args.kernel == "megakernel" else "flag",
            "stack_lds": args.stack,
            "sample_groups": groups_used,
            "sample_groups_chosen_by": "flag" if args.groups else "autotune",
            "bvh": args.bvh,
        },
        "mrays_per_s": rays / wall / 1e6,
        "rays_per_path": rays / max(paths, 1),
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "traffic_source": TRAFFIC_PROFILE_NAME if traffic is not None else None,
            # the profiled traffic over this run's kernel time, against 8 TB/s
            "hbm_gbs": traffic / (kernel_ms_avg * 1e-3) / 1e9 if traffic is not None else None,
            "hbm_frac": traffic / (kernel_ms_avg * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic is not None else None,
            "peak_nonpacked_fp32": FP32_PEAK_NONPACKED_TFLOPS,
            "frac_nonpacked_fp32": achieved_tflops / FP32_PEAK_NONPACKED_TFLOPS,
            "kernel": kname,
            "kernel_ms_avg": kernel_ms_avg,
            "algorithmic_flops_per_query": flops_per_query,
            "work_per_query": {k2: v / queries for k2, v in work.items()},
            # issue-side view from the same committed counter profile:
            # VALU busy = SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE per XCD,
            # lane utilisation = SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)
            "valu_busy": (counter_profile(kname).get("valu_busy")
                          if args.workload == "instanced" else None),
            "valu_lane_utilization": (counter_profile(kname).get("valu_lane_utilization")
                                      if args.workload == "instanced" else None),
        },
        "setup_s": setup_s,
        "bvh_build_ms": {"host_sah_total_build": warm_build_ms, "first_session_with_runtime_start": st0["build_ms"],
                         "gpu_rebuild": st0["bvh_gpu_ms"]},
        # SAH cost of the tree the timed passes traversed (node visits + primitive
        # tests for a ray that hits the root box; sundog_amd/bvhinfo.py)
        "bvh_sah": bvh_sah,
        "finite": finite,
    }

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        flat = scene.flatten()
        host = host_cores()
        threads = args.cpu_threads or host["share"]
        _, _, _, cst = oracle.render(flat, W, H, 0, args.cpu_spp, tile_stride=args.cpu_tile_stride,
                                     threads=threads)
        tiles = ((W + 15) // 16) * ((H + 15) // 16)
        out["cpu_baseline"] = {
            "value": cst["paths"] / cst["seconds"],
            "unit": "paths/s",
            "cores": cst["threads"],
            "kind": "port",
            "nproc": host["nproc"],
            "affinity_cores": host["affinity"],
            "omp_num_threads": host["omp"],
            "cpu_model": host["model"],
            "threads_rule": "the job's CPU share: OMP_NUM_THREADS (16 per GPU on this pool) capped by the affinity mask",
            "sample": "every %dth 16x16 tile of the same %dx%d frame (%d of %d tiles), samples 0..%d; "
                      "%.1f s; %.2f Mrays/s" % (args.cpu_tile_stride, W, H, (tiles + args.cpu_tile_stride - 1)
                                               // args.cpu_tile_stride, tiles, args.cpu_spp - 1,
                                               cst["seconds"],
                                               (cst["radiance_rays"] + cst["shadow_rays"]) / cst["seconds"] / 1e6),
        }
    if args.serving_requests > 0:
        out["serving"] = serving_run(args)
    if not args.no_reference_scenes:
        from sundog_amd import gallery
        out["reference_scenes"] = []
        for name in sorted(gallery.PUBLISHED):
            print("reference scene %s: 1920x1080, 64 spp" % name, flush=True)
            out["reference_scenes"].append(gallery.measure(gallery.program_path(name)))
    print(json.dumps(out), flush=True)


def host_cores():
    """The host's core counts and model, and this job's CPU share."""
    nproc = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    share = aff
    if omp and omp.isdigit() and int(omp) > 0:
        share = min(aff, int(omp))
    return {"nproc": nproc, "affinity": aff, "omp": omp, "model": model, "share": share}


def serving_run(args):
    """The serving workload at a short request count: four resident scenes
    (Cornell; instanced icospheres with 4,096 and 16,384 instances; volumes +
    environment), seeded Poisson arrivals, resolutions {256^2, 512^2,
    1024x576, 1920x1080}, spp log-uniform in [16, 1024]; the same requests
    then rendered one at a time by the per-request megakernel."""
    from sundog_amd import serving
    scenes = serving.build_scenes(args.asset_dir)
    handles = [s.build() for s in scenes]
    srv = serving.Server(budget_bytes=64 << 30, pass_spp=16, tile=16)
    sids = [srv.add_scene(h) for h in handles]
    srv.tune("graph_passes", 4)
    tuned = srv.autotune(sids[1], 1920, 1080, 32)
    reqs = serving.make_requests(args.serving_requests, args.serving_qps, seed=1)
    total = sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    print("serving: %d requests, %.3g paths" % (len(reqs), total), flush=True)
    done, makespan = serving.run(reqs, srv, sids)
    lat = [d["t_done_ms"] - d["t_submit_ms"] for d in done if d["status"] == 2]
    ser_ms = 0.0
    for r in reqs:
        ms, _ = srv.render_serial(sids[r["scene"]], r["width"], r["height"], r["spp"], r["seed"])
        ser_ms += ms
    served = total / makespan
    serial = total / (ser_ms * 1e-3)
    out = {"requests": len(reqs), "completed": len(lat), "qps": args.serving_qps, "paths": total,
           "makespan_s": makespan, "served_paths_per_s": served, "serial_paths_per_s": serial,
           "speedup_vs_serial": served / serial, "p50_ms": serving.percentile(lat, 50),
           "p99_ms": serving.percentile(lat, 99),
           "launch_shape": {k: tuned[k] for k in ("tile", "waves", "stack", "blocks_per_cu", "graph_passes")
                            if k in tuned}}
    if args.serving_qps_sustained > 0:
        # the same requests (same seed: same sizes, scenes and seeds, arrival
        # times scaled) at a rate the server keeps up with
        reqs2 = serving.make_requests(args.serving_requests, args.serving_qps_sustained, seed=1)
        print("serving: sustained run at %.2f qps" % args.serving_qps_sustained, flush=True)
        done2, makespan2 = serving.run(reqs2, srv, sids)
        lat2 = [d["t_done_ms"] - d["t_submit_ms"] for d in done2 if d["status"] == 2]
        out["sustained"] = {"qps": args.serving_qps_sustained, "completed": len(lat2),
                            "offered_paths_per_s": total / reqs2[-1]["arrival_s"],
                            "makespan_s": makespan2, "p50_ms": serving.percentile(lat2, 50),
                            "p99_ms": serving.percentile(lat2, 99)}
    srv.close()
    return out


if __name__ == "__main__":
    main()
