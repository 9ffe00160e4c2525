// Resident render sessions and the reference entry points sundog_render /
// sundog_probe.
//
// sundog_render follows the reference orchestration (src/capi_render.cpp:55-307):
// overrides -> asset load -> mesh-light CDFs -> acceleration build -> film ->
// chunked progressive passes of 16 spp -> output + stats. Device resources
// are owned by a session, so the same code path also serves callers that
// keep a scene resident across many passes (sundog_session_*).
#include "session.h"
#include "transfer.h"

using namespace sdk;

namespace {

thread_local int t_dummy = 0;

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    setLastError(e.what());
    return SUNDOG_ERROR;
  } catch (...) {
    setLastError("unknown error");
    return SUNDOG_ERROR;
  }
}

void writePfm(const std::string& path, const std::vector<float>& rgba, int w, int h) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
  std::vector<float> row((size_t)w * 3);
  for (int y = h - 1; y >= 0; y--) {  // PFM rows run bottom to top
    for (int x = 0; x < w; x++)
      for (int c = 0; c < 3; c++) row[(size_t)x * 3 + c] = rgba[((size_t)y * w + x) * 4 + c];
    fwrite(row.data(), 4, row.size(), f);
  }
  fclose(f);
}

std::string pfmPath(const std::string& p) {
  if (p.size() >= 4 && p.compare(p.size() - 4, 4, ".pfm") == 0) return p;
  return p + ".pfm";
}

std::string jstr(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') { o += '\\'; o += c; }
    else if ((unsigned char)c < 0x20) { char b[8]; snprintf(b, sizeof b, "\\u%04x", c); o += b; }
    else o += c;
  }
  return o + "\"";
}

}  // namespace

// ---------------------------------------------------------------------------
// extern "C": sessions
// ---------------------------------------------------------------------------

extern "C" sundog_session* sundog_session_create(sundog_scene* h, const sundog_render_options* opt) {
  try {
    if (!h) throw std::runtime_error("session: null scene handle");
    if (sundog_scene_validate(h) != SUNDOG_OK) throw std::runtime_error(sundog_last_error());
    applyOverrides(h->scene, opt);  // same in-place override semantics as sundog_render
    std::unique_ptr<sundog_session> s(new sundog_session());
    s->build(h->scene);
    return s.release();
  } catch (const std::exception& e) {
    setLastError(e.what());
    return nullptr;
  }
}

extern "C" void sundog_session_destroy(sundog_session* s) { delete s; }

extern "C" int sundog_session_reset(sundog_session* s) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    s->reset();
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_render(sundog_session* s, int sample_offset, int count) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    if (count < 0) throw std::runtime_error("render: negative sample count");
    s->renderPass(sample_offset, count);
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_set_kernel(sundog_session* s, int kernel) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    if (kernel != SUNDOG_KERNEL_MEGA && kernel != SUNDOG_KERNEL_WAVEFRONT)
      throw std::runtime_error("session: unknown kernel variant");
    s->kernel = kernel;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_set_count_rays(sundog_session* s, int on) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    s->P.countRays = on ? 1 : 0;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_set_tuning(sundog_session* s, const char* key, int value) {
  return guarded([&] {
    if (!s || !key) throw std::runtime_error("session: null argument");
    std::string k = key;
    if (k == "waves") {
      if (value != 0 && (value < 2 || value > 5))
        throw std::runtime_error("tuning: waves must be 0 (auto), 2, 3, 4 or 5");
      s->waves = value;
      if (value == 0) s->resetAuto();
    } else if (k == "stack") {
      if (value < 4 || value > stackLdsMax()) throw std::runtime_error("tuning: stack must be 4.." + std::to_string(stackLdsMax()));
      HIP_OK(hipStreamSynchronize(s->stream));
      s->P.stackLds = value;
    } else if (k == "bvh") {
      if (value < 1 || value > 4)
        throw std::runtime_error("tuning: bvh must be 1 (rebuild both levels with the GPU LBVH), 2 (GPU LBVH "
                                 "mesh levels, host SAH top level over GPU instance boxes), 3 (both levels "
                                 "with GPU PLOC clustering) or 4 (GPU PLOC mesh levels, host SAH top level)");
      s->buildBvhGpu(value == 2 || value == 4, value >= 3 ? 1 : 0);
    } else if (k == "ploc_radius") {
      if (value < 1 || value > 32) throw std::runtime_error("tuning: ploc_radius must be 1..32");
      s->plocRadius = value;
    } else if (k == "wf_waves") {
      if (value < 4 || value > 5) throw std::runtime_error("tuning: wf_waves must be 4 or 5");
      s->wfWaves = value;
    } else if (k == "groups") {
      if (value < 0 || value > kPoolSamples)
        throw std::runtime_error("tuning: groups must be 0 (auto) or 1.." + std::to_string(kPoolSamples));
      s->sampleGroups = value;
      if (value == 0) s->groupPick.reset();
    } else {
      throw std::runtime_error("tuning: unknown key '" + k + "'");
    }
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_autotune(sundog_session* s, int count, int* waves_out) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    if (count <= 0) throw std::runtime_error("autotune: count must be positive");
    if (s->kernel == SUNDOG_KERNEL_WAVEFRONT) throw std::runtime_error("autotune: times the megakernel");
    const bool counting = s->countWork;
    s->countWork = false;
    // a variant fixed by the "waves" tuning is kept; the rest is timed:
    // every candidate twice (the faster pass counts), one untimed warm-up
    // pass and eight timed ones for the occupancy variant, then up to
    // fourteen for the sample groups at that variant, then the occupancy variant
    // again at those groups; later passes of this session use the winners
    const bool tuneWaves = s->waves == 0;
    const bool tuneGroups = s->sampleGroups == 0 && count <= kPoolSamples;
    s->autoPick = OccupancyPicker{};
    s->autoPick.rounds = 2;
    s->autoPick.exhaustive = true;
    s->groupPick = GroupPicker{};
    s->groupPick.rounds = 2;
    s->groupPick.exhaustive = true;
    for (int i = 0; i < 32 && ((tuneWaves && s->autoPick.chosen == 0) || (tuneGroups && s->groupPick.chosen == 0));
         i++)
      s->launchPass(0, count);  // whole passes: the autotuner times full-size launches
    if (tuneWaves && tuneGroups && s->groupPick.chosen != 0) {
      // the occupancy variant again, at the chosen groups (the groups change
      // how many blocks are resident at once, so the best variant can move)
      s->sampleGroups = s->groupPick.chosen;
      s->autoPick = OccupancyPicker{};
      s->autoPick.rounds = 2;
      s->autoPick.exhaustive = true;
      for (int i = 0; i < 9 && s->autoPick.chosen == 0; i++) s->launchPass(0, count);
      s->sampleGroups = 0;  // auto mode, with the picker's choice kept
    }
    s->countWork = counting;
    s->reset();
    if (waves_out) *waves_out = s->currentWaves();
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_set_count_work(sundog_session* s, int on) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    s->countWork = on != 0;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_sync(sundog_session* s) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    HIP_OK(hipStreamSynchronize(s->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_download(sundog_session* s, float* beauty, float* albedo, float* normal) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    size_t n = (size_t)s->P.width * s->P.height * 16;
    if (beauty) HIP_OK(hipMemcpyAsync(beauty, s->accum.p, n, hipMemcpyDeviceToHost, s->stream));
    if (albedo) HIP_OK(hipMemcpyAsync(albedo, s->albedo.p, n, hipMemcpyDeviceToHost, s->stream));
    if (normal) HIP_OK(hipMemcpyAsync(normal, s->normal.p, n, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_bvh(sundog_session* s, sundog_bvh_layout* out, void* nodes, size_t nodesBytes,
                                  float* tris, size_t trisBytes, int32_t* blasRoots, int blasRootsCap) {
  return guarded([&] {
    if (!s || !out) throw std::runtime_error("session: null argument");
    HIP_OK(hipStreamSynchronize(s->stream));
    *out = sundog_bvh_layout{};
    out->num_nodes = s->tlasNodes + s->blasNodes;
    out->tlas_nodes = s->tlasNodes;
    out->num_tris = (int32_t)(s->tris.bytes / (3 * sizeof(float4)));
    out->num_instances = s->numInstances;
    out->tlas_root = s->P.tlasRoot;
    if (nodes) {
      const size_t need = (size_t)out->num_nodes * sizeof(BvhNode4);
      if (nodesBytes < need) throw std::runtime_error("bvh: node buffer needs " + std::to_string(need) + " bytes");
      if (need) HIP_OK(hipMemcpy(nodes, s->nodes.p, need, hipMemcpyDeviceToHost));
    }
    if (tris) {
      const size_t need = (size_t)out->num_tris * 3 * sizeof(float4);
      if (trisBytes < need) throw std::runtime_error("bvh: triangle buffer needs " + std::to_string(need) + " bytes");
      if (need) HIP_OK(hipMemcpy(tris, s->tris.p, need, hipMemcpyDeviceToHost));
    }
    if (blasRoots) {
      for (int i = 0; i < blasRootsCap && i < (int)s->instHost.size(); i++)
        blasRoots[i] = s->instHost[i].geomKind == GK_MESH ? (int32_t)s->instHost[i].blasRoot : -1;
    }
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_get_stats(sundog_session* s, sundog_session_stats* out) {
  return guarded([&] {
    if (!s || !out) throw std::runtime_error("session: null argument");
    HIP_OK(hipStreamSynchronize(s->stream));
    unsigned long long rays = 0, work[4] = {0, 0, 0, 0};
    HIP_OK(hipMemcpy(&rays, s->rayCounter.p, 8, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(work, s->workCounters.p, 32, hipMemcpyDeviceToHost));
    *out = sundog_session_stats{};
    out->rays_traced = rays;
    out->node_visits = work[0];
    out->tri_tests = work[1];
    out->quad_tests = work[2];
    out->inst_transforms = work[3];
    out->paths = s->paths;
    out->render_ms = s->renderMs;
    out->build_ms = s->buildMs;
    out->device_bytes = s->deviceBytes();
    out->stack_depth = s->stackNeed;
    out->tlas_nodes = s->tlasNodes;
    out->bvh_builder = s->bvhBuilder;
    out->bvh_gpu_ms = s->bvhGpuMs;
    out->blas_nodes = s->blasNodes;
    out->waves = s->currentWaves();
    out->stack_lds = s->P.stackLds;
    out->sample_groups = s->currentGroups();
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_size(sundog_session* s, int* w, int* h, int* spp) {
  return guarded([&] {
    if (!s) throw std::runtime_error("session: null handle");
    if (w) *w = s->P.width;
    if (h) *h = s->P.height;
    if (spp) *spp = s->P.sppTotal;
    return SUNDOG_OK;
  });
}

extern "C" void* sundog_session_stream(sundog_session* s) { return s ? (void*)s->stream : nullptr; }

extern "C" int sundog_session_trace(sundog_session* s, const float* rays, int n, int brute, float* t_out,
                                    int32_t* inst_out, int32_t* prim_out) {
  return guarded([&] {
    if (!s || !rays || !t_out || !inst_out || !prim_out) throw std::runtime_error("trace: null argument");
    if (n <= 0) return SUNDOG_OK;
    DevBuf dr, dt, di, dp;
    dr.alloc((size_t)n * 32);
    dt.alloc((size_t)n * 4);
    di.alloc((size_t)n * 4);
    dp.alloc((size_t)n * 4);
    HIP_OK(hipMemcpyAsync(dr.p, rays, (size_t)n * 32, hipMemcpyHostToDevice, s->stream));
    s->ensureSpill(((size_t)n + 255) / 256 * 256);
    HIP_OK(launchTrace(s->P, dr.as<float>(), n, brute != 0, s->numInstances, dt.as<float>(), di.as<int>(),
                       dp.as<int>(), nullptr, s->stream));
    HIP_OK(hipMemcpyAsync(t_out, dt.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipMemcpyAsync(inst_out, di.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipMemcpyAsync(prim_out, dp.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return SUNDOG_OK;
  });
}

// ---------------------------------------------------------------------------
// allocation-free closest-hit query
// ---------------------------------------------------------------------------

namespace {

// Overflow region of the traversal stacks for sundog_session_trace_noalloc:
// a static array on every GPU the library runs on, held until the process
// ends. All calls on one GPU share it (and leave their contents in it), so
// calls that overlap on one GPU race. Nothing in it needs to start at zero:
// a stack entry is always written before it is read.
constexpr size_t kTraceScratchBytes = SUNDOG_TRACE_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kTraceScratchBytes];

constexpr int kMaxGpus = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxGpus] = {};

// Device address of GPU dev's scratch array, looked up once per GPU with
// that GPU current. In host code the array's name is only a stand-in, never
// a device address, so the address from this lookup is what kernels get.
void* traceScratch(int dev) {
  if (dev < 0 || dev >= kMaxGpus) throw std::runtime_error("trace_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    int prev = 0;
    HIP_OK(hipGetDevice(&prev));
    if (prev != dev) HIP_OK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != dev) HIP_OK(hipSetDevice(prev));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("trace_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratchPtr[dev] = ptr;
  }
  return g_scratchPtr[dev];
}

}  // namespace

extern "C" int sundog_session_trace_noalloc(sundog_session* s, const float* d_rays, int n, float* d_t,
                                            int32_t* d_inst, int32_t* d_prim) {
  return guarded([&] {
    if (!s || !d_rays || !d_t || !d_inst || !d_prim) throw std::runtime_error("trace_noalloc: null argument");
    if (n <= 0) return SUNDOG_OK;
    // the query kernel keeps stackLdsEntries() entries per lane in LDS;
    // deeper entries of this scene's trees go to the scratch array, one
    // column per thread of the grid
    const size_t threads = ((size_t)n + 255) / 256 * 256;
    const int rows = s->spillRowsFor(stackLdsEntries());
    const size_t need = (size_t)rows * threads * kStackEntryBytes;
    if (need > kTraceScratchBytes) {
      const size_t maxRays = kTraceScratchBytes / ((size_t)rows * kStackEntryBytes) / 256 * 256;
      char msg[256];
      snprintf(msg, sizeof msg,
               "trace_noalloc: %d rays need %zu bytes of traversal-stack scratch, more than the %zu bytes "
               "available; at most %zu rays per call on this scene",
               n, need, kTraceScratchBytes, maxRays);
      throw std::runtime_error(msg);
    }
    int dev = 0;
    HIP_OK(hipStreamGetDevice(s->stream, &dev));
    LaunchParams P = s->P;
    P.spill = static_cast<uint32_t*>(traceScratch(dev));
    P.spillStride = rows > 0 ? (int)threads : 0;
    P.spillRows = rows;
    HIP_OK(launchTrace(P, d_rays, n, false, s->numInstances, d_t, d_inst, d_prim, nullptr, s->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_shadow(sundog_session* s, const float* rays, int n, float* tr_out) {
  return guarded([&] {
    if (!s || !rays || !tr_out) throw std::runtime_error("shadow: null argument");
    if (n <= 0) return SUNDOG_OK;
    DevBuf dr, dt;
    dr.alloc((size_t)n * 32);
    dt.alloc((size_t)n * 12);
    HIP_OK(hipMemcpyAsync(dr.p, rays, (size_t)n * 32, hipMemcpyHostToDevice, s->stream));
    s->ensureSpill(((size_t)n + 255) / 256 * 256);
    HIP_OK(launchShadow(s->P, dr.as<float>(), n, dt.as<float>(), s->stream));
    HIP_OK(hipMemcpyAsync(tr_out, dt.p, (size_t)n * 12, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_env_bins_mode(sundog_session* s, const float* u, int n, int mode, int32_t* row_out,
                                            int32_t* col_out, float* kernel_ms) {
  return guarded([&] {
    if (!s || !u || !row_out || !col_out) throw std::runtime_error("env bins: null argument");
    if (!s->P.hasEnv) throw std::runtime_error("env bins: scene has no environment map");
    if (mode != 0 && mode != 1) throw std::runtime_error("env bins: mode must be 0 (per-lane table) or 1 (wave-cooperative)");
    if (kernel_ms) *kernel_ms = 0.0f;
    if (n <= 0) return SUNDOG_OK;
    for (int i = 0; i < 2 * n; i++)
      if (!(u[i] >= 0.0f && u[i] < 1.0f)) throw std::runtime_error("env bins: u must lie in [0, 1)");
    DevBuf du, dr, dc;
    du.alloc((size_t)n * 8);
    dr.alloc((size_t)n * 4);
    dc.alloc((size_t)n * 4);
    HIP_OK(hipMemcpyAsync(du.p, u, (size_t)n * 8, hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipEventRecord(s->ev0, s->stream));
    HIP_OK(launchEnvBins(s->P, du.as<float>(), n, dr.as<int>(), dc.as<int>(), s->stream, mode));
    HIP_OK(hipEventRecord(s->ev1, s->stream));
    HIP_OK(hipMemcpyAsync(row_out, dr.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipMemcpyAsync(col_out, dc.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    if (kernel_ms) *kernel_ms = ms;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_env_bins(sundog_session* s, const float* u, int n, int32_t* row_out,
                                       int32_t* col_out) {
  return sundog_session_env_bins_mode(s, u, n, 0, row_out, col_out, nullptr);
}

extern "C" int sundog_session_light_samples(sundog_session* s, int light, const float* xs, int n, uint64_t seed0,
                                           uint64_t seq, float* out) {
  return guarded([&] {
    if (!s || !xs || !out) throw std::runtime_error("light samples: null argument");
    if (light < 0 || light >= s->P.numLights) throw std::runtime_error("light samples: light id out of range");
    if (n <= 0) return SUNDOG_OK;
    DevBuf dx, dout;
    dx.alloc((size_t)n * 12);
    dout.alloc((size_t)n * 48);
    HIP_OK(hipMemcpyAsync(dx.p, xs, (size_t)n * 12, hipMemcpyHostToDevice, s->stream));
    HIP_OK(launchLightSample(s->P, light, dx.as<float>(), n, seed0, seq, dout.as<float>(), s->stream));
    HIP_OK(hipMemcpyAsync(out, dout.p, (size_t)n * 48, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_session_flame_field(sundog_session* s, int flame, const float* pts, int n, float* sigma_out,
                                          float* emission_out) {
  return guarded([&] {
    if (!s || !pts || !sigma_out || !emission_out) throw std::runtime_error("flame field: null argument");
    if (flame < 0 || flame >= s->P.numFlames) throw std::runtime_error("flame field: flame id out of range");
    if (n <= 0) return SUNDOG_OK;
    DevBuf dp, ds, de;
    dp.alloc((size_t)n * 12);
    ds.alloc((size_t)n * 4);
    de.alloc((size_t)n * 12);
    HIP_OK(hipMemcpyAsync(dp.p, pts, (size_t)n * 12, hipMemcpyHostToDevice, s->stream));
    HIP_OK(launchFlameField(s->P, flame, dp.as<float>(), n, ds.as<float>(), de.as<float>(), s->stream));
    HIP_OK(hipMemcpyAsync(sigma_out, ds.p, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipMemcpyAsync(emission_out, de.p, (size_t)n * 12, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_device_pcg32(uint64_t seed, uint64_t seq, int n, uint32_t* out) {
  return guarded([&] {
    if (!out || n < 0) throw std::runtime_error("pcg32: bad arguments");
    if (n == 0) return SUNDOG_OK;
    DevBuf d;
    d.alloc((size_t)n * 4);
    HIP_OK(launchPcg(seed, seq, n, d.as<uint32_t>(), nullptr));
    HIP_OK(hipMemcpy(out, d.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    return SUNDOG_OK;
  });
}

// ---------------------------------------------------------------------------
// extern "C": reference entry points
// ---------------------------------------------------------------------------

extern "C" int sundog_probe(sundog_device_info* out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("probe: null output");
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    *out = sundog_device_info{};
    snprintf(out->name, sizeof out->name, "%s (%s)", prop.name, prop.gcnArchName);
    out->cc_major = prop.major;
    out->cc_minor = prop.minor;
    int dv = 0, rv = 0;
    (void)hipDriverGetVersion(&dv);
    (void)hipRuntimeGetVersion(&rv);
    out->driver_version = dv;
    out->runtime_version = rv;
    out->total_mem_mb = prop.totalGlobalMem >> 20;
    out->optix_header_version = 0;
    out->rtcore_version = 0;
    return SUNDOG_OK;
  });
}

namespace sdn {
double denoiseFilms(const char* weightsPath, const float4* beauty, const float4* albedo, const float4* normal,
                    int w, int h, float* outHost);
}

extern "C" int sundog_render(sundog_scene* h, const sundog_render_options* opt) {
  return guarded([&] {
    if (!opt || !opt->out_path || !*opt->out_path) throw std::runtime_error("render: out_path is required");
    if (sundog_scene_validate(h) != SUNDOG_OK) throw std::runtime_error(sundog_last_error());
    if (!std::isnan(opt->physics_time) && opt->physics_time >= 0.0 && !h->scene.physics.enabled)
      throw std::runtime_error("--physics-time: scene has no physics block");
    const char* name = opt->scene_name ? opt->scene_name : "(scene)";
    bool quiet = opt->quiet != 0;
    auto tTotal = std::chrono::steady_clock::now();
    size_t freeBefore = 0, totalMem = 0;
    HIP_OK(hipMemGetInfo(&freeBefore, &totalMem));
    std::unique_ptr<sundog_session> s(sundog_session_create(h, opt));
    if (!s) throw std::runtime_error(sundog_last_error());
    if (h->scene.physics.enabled && !quiet)
      fprintf(stderr, "physics: settling is not part of this build; rendering the scene's baked poses\n");
    const RenderSettings& rs = s->prep.scene.render;
    s->P.countRays = opt->stats_path ? 1 : 0;
    sundog_device_info di{};
    sundog_probe(&di);
    if (!quiet)
      fprintf(stderr, "sundog: %s %dx%d, %d spp, depth %d%s on %s\n", name, rs.width, rs.height, rs.spp,
              rs.maxDepth, rs.denoise ? " [denoise]" : "", di.name);
    auto tRender = std::chrono::steady_clock::now();
    int done = 0;
    while (done < rs.spp) {
      int chunk = std::min(rs.chunk, rs.spp - done);
      s->renderPass(done, chunk);
      done += chunk;
      if (!quiet) {
        fprintf(stderr, "\r%d/%d spp (%.1f%%)", done, rs.spp, 100.0 * done / rs.spp);
        fflush(stderr);
      }
    }
    double renderMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tRender).count();
    if (!quiet) fprintf(stderr, "\n");
    size_t npx = (size_t)rs.width * rs.height;
    std::vector<float> img(npx * 4);
    // Denoise on request (src/capi_render.cpp:247-256): the learned denoiser
    // runs on the device films when trained weights are named by
    // SUNDOG_DENOISER_WEIGHTS; otherwise the plain render is the output.
    bool denoised = false;
    double denoiseMs = 0.0;
    if (rs.denoise) {
      const char* weights = getenv("SUNDOG_DENOISER_WEIGHTS");
      if (weights && *weights) {
        HIP_OK(hipStreamSynchronize(s->stream));
        denoiseMs = sdn::denoiseFilms(weights, s->accum.as<float4>(), s->albedo.as<float4>(),
                                      s->normal.as<float4>(), rs.width, rs.height, img.data());
        denoised = true;
        if (!quiet) fprintf(stderr, "denoise: %.1f ms (%s)\n", denoiseMs, weights);
      } else if (!quiet) {
        fprintf(stderr, "denoise: SUNDOG_DENOISER_WEIGHTS names no trained weights; writing the plain render\n");
      }
    }
    if (!denoised) sundog_session_download(s.get(), img.data(), nullptr, nullptr);
    std::string out = pfmPath(opt->out_path);
    writePfm(out, img, rs.width, rs.height);
    // display-referred codes of the reference's AVIF (PQ / BT.2020, 12-bit)
    std::vector<uint16_t> pq(npx * 3);
    encodePq12(img.data(), rs.width, rs.height, rs.exposure, pq.data());
    writePpm12(out.substr(0, out.size() - 4) + ".pq.ppm", pq.data(), rs.width, rs.height);
    if (opt->aov_albedo_path) {
      sundog_session_download(s.get(), nullptr, img.data(), nullptr);
      writePfm(pfmPath(opt->aov_albedo_path), img, rs.width, rs.height);
    }
    if (opt->aov_normal_path) {
      sundog_session_download(s.get(), nullptr, nullptr, img.data());
      writePfm(pfmPath(opt->aov_normal_path), img, rs.width, rs.height);
    }
    size_t freeAfter = 0;
    HIP_OK(hipMemGetInfo(&freeAfter, &totalMem));
    sundog_session_stats st{};
    sundog_session_get_stats(s.get(), &st);
    double totalMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tTotal).count();
    if (!quiet) {
      fprintf(stderr, "Rendering elapsed time: %f seconds", renderMs / 1000.0);
      if (st.rays_traced) fprintf(stderr, " (%.0f Mrays/s)", st.rays_traced / renderMs / 1000.0);
      fprintf(stderr, "\nwrote %s\n", out.c_str());
    }
    if (opt->stats_path) {
      FILE* f = fopen(opt->stats_path, "w");
      if (!f) throw std::runtime_error(std::string("cannot write ") + opt->stats_path);
      const Scene& S = s->prep.scene;
      size_t tris = 0;
      for (const HostMesh& m : s->prep.meshes) tris += m.numTris();
      fprintf(f, "{\n  \"scene\": %s,\n", jstr(name).c_str());
      fprintf(f, "  \"width\": %d,\n  \"height\": %d,\n  \"spp\": %d,\n", rs.width, rs.height, rs.spp);
      fprintf(f, "  \"max_depth\": %d,\n  \"seed\": %u,\n", rs.maxDepth, rs.seed);
      fprintf(f, "  \"denoised\": %s,\n", denoised ? "true" : "false");
      fprintf(f,
              "  \"timings_ms\": {\n    \"scene_load\": %.17g,\n    \"physics\": 0,\n"
              "    \"gas_build\": %.17g,\n    \"render\": %.17g,\n    \"denoise\": %.17g,\n"
              "    \"total\": %.17g\n  },\n",
              s->prep.loadMs, s->buildMs - s->prep.loadMs, renderMs, denoiseMs, totalMs);
      fprintf(f, "  \"rays_traced\": %llu,\n  \"mrays_per_sec\": %.17g,\n", (unsigned long long)st.rays_traced,
              st.rays_traced / renderMs / 1000.0);
      fprintf(f, "  \"peak_vram_mb\": %zu,\n", freeBefore > freeAfter ? (freeBefore - freeAfter) >> 20 : (size_t)0);
      fprintf(f,
              "  \"scene_stats\": {\n    \"objects\": %d,\n    \"lights\": %d,\n    \"meshes\": %d,\n"
              "    \"mesh_triangles\": %zu\n  },\n",
              (int)S.objects.size(), (int)S.lights.size(), (int)S.meshes.size(), tris);
      fprintf(f,
              "  \"device\": {\n    \"name\": %s,\n    \"compute\": \"gfx%d%d\",\n    \"vram_mb\": %llu,\n"
              "    \"hip_driver\": %d,\n    \"hip_runtime\": %d,\n    \"optix_header\": 0\n  }\n}\n",
              jstr(di.name).c_str(), di.cc_major, di.cc_minor, (unsigned long long)di.total_mem_mb,
              di.driver_version, di.runtime_version);
      fclose(f);
    }
    (void)t_dummy;
    return SUNDOG_OK;
  });
}
