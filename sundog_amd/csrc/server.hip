// Serving loop: resident scenes, admission by HBM footprint, and continuous
// batching of render requests through the persistent tile-queue kernel.
//
// The reference renders one scene per call and frees every device resource
// on the way out (src/capi_render.cpp:65, :304); it has no request loop.
// Here a worker thread owns the GPU stream. Each batch renders the next pass
// (pass_spp samples) of every active request in ONE launch of the persistent
// kernel, whose waves pull (request, tile) items from a shared counter.
// Requests that arrive while a batch runs are admitted into the next batch
// as soon as their film fits the HBM budget; requests that finish leave the
// batch and their film is copied to host memory and freed. Per-pixel sample
// order is the same as a render of the request alone, so each result is
// bit-identical alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "session.h"

using namespace sdk;

namespace {

using Clock = std::chrono::steady_clock;

template <class F>
int guardedS(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    setLastError(e.what());
    return SUNDOG_ERROR;
  }
}

}  // namespace

enum { REQ_QUEUED = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_REJECTED = 3, REQ_FAILED = 4 };

struct ServerRequest {
  int64_t id = 0;
  int scene = 0;
  int width = 0, height = 0, spp = 0;
  uint32_t seed = 0;
  int done = 0;  // samples completed
  int status = REQ_QUEUED;
  size_t bytes = 0;
  double tSubmit = 0, tAdmit = 0, tDone = 0;
  DevBuf accum, albedo, normal;
  std::vector<float> beauty, alb, nrm;  // host copies once done
  LaunchParams P{};
};

// Largest group buffer of a batch, in 64-pixel pool regions (36 KiB each:
// 9 planes x 16 samples x 64 pixels of FP32): batches whose tiles hold more
// pools are not split into sample groups.
constexpr int64_t kMaxGroupPools = 16384;

struct sundog_server {
  std::vector<std::unique_ptr<sundog_session>> scenes;
  size_t budget = 0;
  size_t used = 0;
  int passSpp = 16;
  int tileSize = 16;
  int waves = 4;         // occupancy variant of the persistent kernel
  int stackLds = 16;     // LDS traversal-stack entries per lane
  int blocksReq = 0;     // queue depth: resident persistent blocks per CU (0 = as many as fit)
  int graphPasses = 4;   // progressive passes captured into one graph launch (1 = no graph)
  // sample groups per pass: 0 splits each tile's pass over enough work items
  // to reach the resident waves when a batch's tiles alone cannot (at most
  // pass_spp groups, and only with pass_spp <= 16); 1 never splits; 2..16
  // always split that many ways (capped at pass_spp)
  int groupsReq = 0;
  int serialWaves = 0;   // occupancy variant of the per-request megakernel baseline (0: picked per request)
  int blocksPerCU = 1;
  int numCUs = 1;
  int maxStackNeed = 0;  // deepest traversal stack over the resident scenes
  hipStream_t stream = nullptr;
  DevBuf dReqs, dItems, dQueue, dFirst, spill, sampleBuf, rayCounter, groupBuf;
  int64_t sampleBufWaves = 0;  // per-sample result regions (one per wave of the persistent grid)
  size_t reqCap = 0, itemCap = 0, queueCap = 0;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv, idleCv;
  std::deque<ServerRequest*> waiting;
  std::vector<ServerRequest*> active;
  std::map<int64_t, std::unique_ptr<ServerRequest>> all;
  std::deque<int64_t> completed;  // ids not yet reported by poll
  int64_t nextId = 0;
  bool stop = false;
  bool started = false;
  // a serial render or the autotuner runs on the caller's thread on the
  // server stream: no batch may run meanwhile (guarded by mu)
  bool exclusive = false;
  std::string workerError;
  std::thread worker;
  // stats
  uint64_t paths = 0, batches = 0, items = 0, graphLaunches = 0;
  double kernelMs = 0, busyMs = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  double now() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

  ~sundog_server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (worker.joinable()) worker.join();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  LaunchParams requestParams(const ServerRequest& r) const {
    const sundog_session& sc = *scenes[r.scene];
    LaunchParams P = sc.P;
    P.width = r.width;
    P.height = r.height;
    P.sppTotal = r.spp;
    P.seed = r.seed;
    P.countRays = 1;
    P.accum = r.accum.as<float4>();
    P.albedo = r.albedo.as<float4>();
    P.normal = r.normal.as<float4>();
    P.rayCounter = rayCounter.as<unsigned long long>();
    P.workCounters = nullptr;
    Camera c = makeCamera(sc.prep.scene.camera, r.width, r.height);
    auto st3 = [](float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
    st3(P.cam.origin, c.origin);
    st3(P.cam.lowerLeft, c.lowerLeft);
    st3(P.cam.horizontal, c.horizontal);
    st3(P.cam.vertical, c.vertical);
    st3(P.cam.u, c.u);
    st3(P.cam.v, c.v);
    st3(P.cam.w, c.w);
    P.cam.lensRadius = c.lensRadius;
    return P;
  }

  static size_t filmBytes(int w, int h) { return (size_t)w * h * 48; }

  void ensureDevice() {
    if (stream) return;
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_OK(hipEventCreate(&ev0));
    HIP_OK(hipEventCreate(&ev1));
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    numCUs = prop.multiProcessorCount;
    rayCounter.alloc(8);
    HIP_OK(hipMemsetAsync(rayCounter.p, 0, 8, stream));
    applyTuning();
  }

  // Launch-shape knobs (buffer sizes, tile, pass length, graph passes) may
  // only change while no request is in flight, so the worker never reads
  // one mid-change; the caller holds mu.
  void requireIdle(const char* what) const {
    if (!active.empty() || !waiting.empty())
      throw std::runtime_error(std::string("tuning: ") + what + " cannot change while requests are in flight");
  }

  void applyTuning() {
    const int fit = persistentBlocksPerCU(waves, stackLds);
    blocksPerCU = blocksReq > 0 ? std::min(blocksReq, fit) : fit;
    size_t threads = (size_t)numCUs * blocksPerCU * renderBlockThreads();
    int rows = std::max(0, maxStackNeed - stackLds);
    spill.alloc(std::max<size_t>(16, (size_t)rows * threads * kStackEntryBytes));
    sampleBufWaves = (int64_t)(threads / 64);
    sampleBuf.alloc((size_t)sampleBufWaves * kPoolPlanes * kPoolSamples * 64 * sizeof(float));
  }

  RequestDev requestRecord(const ServerRequest& r, int sceneSlot) const {
    const LaunchParams P = requestParams(r);
    RequestDev d{};
    d.scene = sceneSlot;
    d.width = r.width;
    d.height = r.height;
    d.sppTotal = r.spp;
    d.sampleOffset = r.done;
    d.sppThisLaunch = std::min(passSpp, r.spp - r.done);
    d.passSpp = passSpp;
    d.seed = r.seed;
    d.accum = P.accum;
    d.albedo = P.albedo;
    d.normal = P.normal;
    d.cam = P.cam;
    return d;
  }

  // One batch: the next `passes` progressive passes of every active request.
  // Requests are grouped so that one persistent launch references at most
  // kMaxScenesPerLaunch scenes (their parameters travel as a kernel
  // argument). Each pass is a k_queue_prep (advances every request's sample
  // window from its first sample and the group's pass counter) followed by
  // one k_persistent launch per group; with passes > 1 the whole sequence
  // is captured into a hipGraph and launched once. Returns the number of
  // passes run. Called by the worker with the lock released; only the
  // worker touches `active`'s films.
  int runBatch(std::vector<ServerRequest*>& batch, int passes) {
    std::vector<ServerRequest*> order(batch);
    std::stable_sort(order.begin(), order.end(),
                     [](const ServerRequest* a, const ServerRequest* b) { return a->scene < b->scene; });
    struct Group {
      SceneTable table;
      int numScenes = 0;
      size_t reqBegin = 0, reqEnd = 0, itemBegin = 0, itemEnd = 0;
    };
    std::vector<Group> groups;
    std::vector<RequestDev> reqs;
    std::vector<int32_t> first;
    std::vector<WorkItem> work;
    int lastScene = -1;
    // sample groups of this batch's passes (the same for every request)
    const int poolsPerTile = tileSize * tileSize / 64;
    const int residentWaves = numCUs * blocksPerCU * (renderBlockThreads() / 64);
    int64_t totalTiles = 0;
    for (const ServerRequest* rp : order)
      totalTiles += (int64_t)((rp->width + tileSize - 1) / tileSize) * ((rp->height + tileSize - 1) / tileSize);
    int sampleGroups = 1;
    if (groupsReq != 1 && passSpp >= 2 && passSpp <= kPoolSamples && totalTiles * poolsPerTile <= kMaxGroupPools) {
      if (groupsReq >= 2) sampleGroups = std::min(groupsReq, passSpp);
      else if (totalTiles < residentWaves)
        sampleGroups = (int)std::min<int64_t>(passSpp, (residentWaves + totalTiles - 1) / totalTiles);
    }
    if (sampleGroups > 1) {
      const size_t need = (size_t)(totalTiles * poolsPerTile) * kPoolPlanes * kPoolSamples * 64 * sizeof(float);
      if (groupBuf.bytes < need) groupBuf.alloc(need);
    }
    int64_t regionNext = 0;
    for (ServerRequest* rp : order) {
      ServerRequest& r = *rp;
      if (groups.empty() || (r.scene != lastScene && groups.back().numScenes == kMaxScenesPerLaunch)) {
        groups.emplace_back();
        groups.back().reqBegin = reqs.size();
        groups.back().itemBegin = work.size();
        lastScene = -1;
      }
      Group& g = groups.back();
      if (r.scene != lastScene) {
        LaunchParams P = scenes[r.scene]->P;
        P.countRays = 1;
        P.rayCounter = rayCounter.as<unsigned long long>();
        P.workCounters = nullptr;
        g.table.scene[g.numScenes++] = P;
        lastScene = r.scene;
      }
      int local = (int)(reqs.size() - g.reqBegin);
      const int tx = (r.width + tileSize - 1) / tileSize, ty = (r.height + tileSize - 1) / tileSize;
      RequestDev rd = requestRecord(r, g.numScenes - 1);
      rd.groups = sampleGroups;
      rd.regionBase = (int32_t)regionNext;
      if (sampleGroups > 1) regionNext += (int64_t)tx * ty * poolsPerTile;
      reqs.push_back(rd);
      first.push_back(r.done);
      for (int t = 0; t < tx * ty; t++)
        for (int k = 0; k < sampleGroups; k++) work.push_back(WorkItem{(int32_t)local, t, k, 0});
      g.reqEnd = reqs.size();
      g.itemEnd = work.size();
    }
    if (reqs.size() > reqCap) {
      reqCap = reqs.size() * 2;
      dReqs.alloc(reqCap * sizeof(RequestDev));
      dFirst.alloc(reqCap * sizeof(int32_t));
    }
    if (work.size() > itemCap) {
      itemCap = work.size() * 2;
      dItems.alloc(itemCap * sizeof(WorkItem));
    }
    if (groups.size() > queueCap) {
      queueCap = groups.size() * 2;
      dQueue.alloc(queueCap * sizeof(QueueState));
    }
    HIP_OK(hipMemcpyAsync(dReqs.p, reqs.data(), reqs.size() * sizeof(RequestDev), hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(dFirst.p, first.data(), first.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    HIP_OK(hipMemcpyAsync(dItems.p, work.data(), work.size() * sizeof(WorkItem), hipMemcpyHostToDevice, stream));
    const int blocks = numCUs * blocksPerCU;
    const size_t threads = (size_t)blocks * renderBlockThreads();
    auto enqueue = [&]() {
      HIP_OK(launchQueueReset(dQueue.as<QueueState>(), (int)groups.size(), stream));
      for (int p = 0; p < passes; p++) {
        for (size_t gi = 0; gi < groups.size(); gi++) {
          const Group& g = groups[gi];
          QueueState* q = dQueue.as<QueueState>() + gi;
          RequestDev* rq = dReqs.as<RequestDev>() + g.reqBegin;
          HIP_OK(launchQueuePrep(q, rq, dFirst.as<int32_t>() + g.reqBegin, (int)(g.reqEnd - g.reqBegin), stream));
          const WorkItem* gi0 = dItems.as<WorkItem>() + g.itemBegin;
          const int nItems = (int)(g.itemEnd - g.itemBegin);
          HIP_OK(launchPersistent(g.table, rq, gi0, nItems, q, tileSize, waves, blocks, spill.as<uint32_t>(),
                                  (int)threads, stackLds, sampleBuf.as<float>(), sampleBufWaves,
                                  groupBuf.as<float>(), rayCounter.as<unsigned long long>(), stream));
          if (sampleGroups > 1) HIP_OK(launchFoldGroups(rq, gi0, nItems, tileSize, groupBuf.as<float>(), stream));
        }
      }
    };
    HIP_OK(hipEventRecord(ev0, stream));
    if (passes > 1) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      try {
        enqueue();
      } catch (...) {
        (void)hipStreamEndCapture(stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      HIP_OK(hipStreamEndCapture(stream, &graph));
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e == hipSuccess) e = hipGraphLaunch(exec, stream);
      if (exec) (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      HIP_OK(e);
      graphLaunches++;
    } else {
      enqueue();
    }
    HIP_OK(hipEventRecord(ev1, stream));
    HIP_OK(hipEventSynchronize(ev1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, ev0, ev1));
    kernelMs += ms;
    batches++;
    items += work.size() * passes;
    for (ServerRequest* rp : batch) {
      ServerRequest& r = *rp;
      int n = std::min(passes * passSpp, r.spp - r.done);
      paths += (uint64_t)r.width * r.height * n;
      r.done += n;
    }
    return passes;
  }

  // Passes for the next batch: as many as every active request still needs
  // (so no request idles inside a graph), capped by graphPasses.
  int passesFor(const std::vector<ServerRequest*>& batch) const {
    int p = graphPasses;
    for (const ServerRequest* r : batch) p = std::min(p, (r->spp - r->done + passSpp - 1) / passSpp);
    return std::max(1, p);
  }

  void finish(ServerRequest& r) {
    size_t n = (size_t)r.width * r.height * 4;
    r.beauty.resize(n);
    r.alb.resize(n);
    r.nrm.resize(n);
    HIP_OK(hipMemcpyAsync(r.beauty.data(), r.accum.p, n * 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(r.alb.data(), r.albedo.p, n * 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipMemcpyAsync(r.nrm.data(), r.normal.p, n * 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    r.accum.release();
    r.albedo.release();
    r.normal.release();
  }

  void admitLocked() {
    while (!waiting.empty()) {
      ServerRequest* r = waiting.front();
      if (used + r->bytes > budget) break;  // FIFO: wait for memory to free up
      waiting.pop_front();
      r->tAdmit = now();
      r->status = REQ_RUNNING;
      used += r->bytes;
      active.push_back(r);
    }
  }

  void loop() {
    try {
      while (true) {
        std::vector<ServerRequest*> batch;
        {
          std::unique_lock<std::mutex> lk(mu);
          admitLocked();
          while (active.empty() && !stop) {
            idleCv.notify_all();
            cv.wait(lk);
            admitLocked();
          }
          if (stop && active.empty()) break;
          batch = active;
        }
        // films for newly admitted requests
        for (ServerRequest* r : batch) {
          if (!r->accum.p) {
            size_t n = (size_t)r->width * r->height * 16;
            r->accum.alloc(n);
            r->albedo.alloc(n);
            r->normal.alloc(n);
            HIP_OK(hipMemsetAsync(r->accum.p, 0, n, stream));
            HIP_OK(hipMemsetAsync(r->albedo.p, 0, n, stream));
            HIP_OK(hipMemsetAsync(r->normal.p, 0, n, stream));
          }
        }
        double tb = now();
        runBatch(batch, passesFor(batch));
        std::vector<ServerRequest*> finished;
        for (ServerRequest* r : batch)
          if (r->done >= r->spp) finished.push_back(r);
        for (ServerRequest* r : finished) finish(*r);
        busyMs += now() - tb;
        {
          std::lock_guard<std::mutex> lk(mu);
          for (ServerRequest* r : finished) {
            r->tDone = now();
            r->status = REQ_DONE;
            used -= r->bytes;
            active.erase(std::find(active.begin(), active.end(), r));
            completed.push_back(r->id);
          }
        }
      }
    } catch (const std::exception& e) {
      // let whatever was queued on the stream end before its films go
      (void)hipStreamSynchronize(stream);
      std::lock_guard<std::mutex> lk(mu);
      workerError = e.what();
      auto failOne = [&](ServerRequest* r) {
        r->status = REQ_FAILED;
        r->tDone = now();
        r->accum.release();
        r->albedo.release();
        r->normal.release();
        completed.push_back(r->id);
      };
      for (ServerRequest* r : active) failOne(r);
      for (ServerRequest* r : waiting) failOne(r);
      active.clear();
      waiting.clear();
      used = 0;
    }
    idleCv.notify_all();
  }

  void start() {
    if (started) return;
    ensureDevice();
    started = true;
    worker = std::thread([this] { loop(); });
  }
};

// Claims the server stream for a call that drives it from the caller's
// thread (serial render, autotuner): refused while requests are queued or
// running, and submits are refused until it ends. Snapshots the tuning the
// call reads.
struct ExclusiveUse {
  sundog_server& s;
  int serialWaves = 0;
  ExclusiveUse(sundog_server& srv, const char* who) : s(srv) {
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.active.empty() || !s.waiting.empty() || s.exclusive)
      throw std::runtime_error(std::string(who) + ": the server is busy");
    s.exclusive = true;
    serialWaves = s.serialWaves;
  }
  ~ExclusiveUse() {
    std::lock_guard<std::mutex> lk(s.mu);
    s.exclusive = false;
  }
};

extern "C" sundog_server* sundog_server_create(uint64_t hbm_budget_bytes, int pass_spp, int tile_size) {
  try {
    if (pass_spp <= 0) pass_spp = 16;
    if (tile_size != 8 && tile_size != 16 && tile_size != 32)
      throw std::runtime_error("server: tile size must be 8, 16 or 32");
    std::unique_ptr<sundog_server> s(new sundog_server());
    s->budget = hbm_budget_bytes;
    s->passSpp = pass_spp;
    s->tileSize = tile_size;
    s->t0 = Clock::now();
    return s.release();
  } catch (const std::exception& e) {
    setLastError(e.what());
    return nullptr;
  }
}

extern "C" void sundog_server_destroy(sundog_server* s) { delete s; }

extern "C" int sundog_server_add_scene(sundog_server* s, sundog_scene* h, const sundog_render_options* opt) {
  try {
    if (!s || !h) throw std::runtime_error("server: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) throw std::runtime_error("server: scenes must be added before the first submit");
    sundog_session* sc = sundog_session_create(h, opt);
    if (!sc) throw std::runtime_error(sundog_last_error());
    s->scenes.emplace_back(sc);
    s->maxStackNeed = std::max(s->maxStackNeed, sc->stackNeed);
    if (s->stream) s->applyTuning();
    return (int)s->scenes.size() - 1;
  } catch (const std::exception& e) {
    setLastError(e.what());
    return -1;
  }
}

extern "C" int sundog_server_set_tuning(sundog_server* s, const char* key, int value) {
  return guardedS([&] {
    if (!s || !key) throw std::runtime_error("server: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::string k = key;
    if (s->exclusive) throw std::runtime_error("tuning: a serial render or the autotuner holds the server");
    if (k == "tile") {
      if (value != 8 && value != 16 && value != 32) throw std::runtime_error("tuning: tile must be 8, 16 or 32");
      s->requireIdle("tile");
      s->tileSize = value;
    } else if (k == "waves") {
      if (value < 2 || value > 5) throw std::runtime_error("tuning: waves must be 2, 3, 4 or 5");
      s->requireIdle("waves");
      s->waves = value;
      if (s->stream) s->applyTuning();
    } else if (k == "stack") {
      if (value < 4 || value > stackLdsMax()) throw std::runtime_error("tuning: stack must be 4.." + std::to_string(stackLdsMax()));
      s->requireIdle("stack");
      s->stackLds = value;
      if (s->stream) s->applyTuning();
    } else if (k == "blocks") {
      if (value < 0) throw std::runtime_error("tuning: blocks must be >= 0 (0 = as many as fit)");
      s->requireIdle("blocks");
      s->blocksReq = value;
      if (s->stream) s->applyTuning();
    } else if (k == "graph_passes") {
      if (value < 1 || value > 64) throw std::runtime_error("tuning: graph_passes must be 1..64");
      s->requireIdle("graph_passes");
      s->graphPasses = value;
    } else if (k == "serial_waves") {
      if (value != 0 && (value < 2 || value > 5))
        throw std::runtime_error("tuning: serial_waves must be 0 (auto) or 2..5");
      s->serialWaves = value;
    } else if (k == "groups") {
      if (value < 0 || value > kPoolSamples)
        throw std::runtime_error("tuning: groups must be 0 (auto), 1 (never split) or 2.." + std::to_string(kPoolSamples));
      s->requireIdle("groups");
      s->groupsReq = value;
    } else if (k == "pass_spp") {
      if (value <= 0) throw std::runtime_error("tuning: pass_spp must be positive");
      s->requireIdle("pass_spp");
      s->passSpp = value;
    } else {
      throw std::runtime_error("tuning: unknown key '" + k + "'");
    }
    return SUNDOG_OK;
  });
}

extern "C" int64_t sundog_server_submit(sundog_server* s, int scene_id, int width, int height, int spp,
                                        int64_t seed) {
  try {
    if (!s) throw std::runtime_error("server: null handle");
    if (scene_id < 0 || scene_id >= (int)s->scenes.size()) throw std::runtime_error("submit: scene id out of range");
    if (width <= 0 || height <= 0 || spp <= 0) throw std::runtime_error("submit: width, height and spp must be positive");
    std::unique_ptr<ServerRequest> r(new ServerRequest());
    r->scene = scene_id;
    r->width = width;
    r->height = height;
    r->spp = spp;
    r->seed = seed >= 0 ? (uint32_t)seed : s->scenes[scene_id]->P.seed;
    r->bytes = sundog_server::filmBytes(width, height);
    int64_t id;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (!s->workerError.empty()) throw std::runtime_error("server failed: " + s->workerError);
      if (s->exclusive) throw std::runtime_error("submit: a serial render or the autotuner holds the server");
      id = s->nextId++;
      r->id = id;
      r->tSubmit = s->now();
      ServerRequest* raw = r.get();
      s->all[id] = std::move(r);
      if (raw->bytes > s->budget) {
        raw->status = REQ_REJECTED;
        raw->tDone = raw->tSubmit;
        s->completed.push_back(id);
      } else {
        s->waiting.push_back(raw);
      }
    }
    s->start();
    s->cv.notify_all();
    return id;
  } catch (const std::exception& e) {
    setLastError(e.what());
    return -1;
  }
}

extern "C" int sundog_server_poll(sundog_server* s, sundog_request_info* out, int max) {
  if (!s || (!out && max > 0)) {
    setLastError("poll: null argument");
    return -1;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  int n = 0;
  while (n < max && !s->completed.empty()) {
    int64_t id = s->completed.front();
    s->completed.pop_front();
    const ServerRequest& r = *s->all[id];
    sundog_request_info& i = out[n++];
    i = sundog_request_info{};
    i.id = id;
    i.status = r.status;
    i.scene = r.scene;
    i.width = r.width;
    i.height = r.height;
    i.spp = r.spp;
    i.t_submit_ms = r.tSubmit;
    i.t_admit_ms = r.tAdmit;
    i.t_done_ms = r.tDone;
    i.paths = (uint64_t)r.width * r.height * r.done;
  }
  return n;
}

extern "C" int sundog_server_fetch(sundog_server* s, int64_t id, float* beauty, float* albedo, float* normal) {
  return guardedS([&] {
    if (!s) throw std::runtime_error("fetch: null handle");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) throw std::runtime_error("fetch: unknown request id");
    ServerRequest& r = *it->second;
    if (r.status != REQ_DONE) throw std::runtime_error("fetch: request is not complete");
    size_t n = (size_t)r.width * r.height * 4;
    if (beauty) memcpy(beauty, r.beauty.data(), n * 4);
    if (albedo) memcpy(albedo, r.alb.data(), n * 4);
    if (normal) memcpy(normal, r.nrm.data(), n * 4);
    return SUNDOG_OK;
  });
}

extern "C" int sundog_server_release(sundog_server* s, int64_t id) {
  return guardedS([&] {
    if (!s) throw std::runtime_error("release: null handle");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) throw std::runtime_error("release: unknown request id");
    if (it->second->status == REQ_QUEUED || it->second->status == REQ_RUNNING)
      throw std::runtime_error("release: request still in flight");
    s->all.erase(it);
    return SUNDOG_OK;
  });
}

extern "C" int sundog_server_wait_idle(sundog_server* s, double timeout_s) {
  return guardedS([&] {
    if (!s) throw std::runtime_error("wait: null handle");
    std::unique_lock<std::mutex> lk(s->mu);
    auto idle = [&] { return (s->active.empty() && s->waiting.empty()) || !s->workerError.empty(); };
    if (timeout_s <= 0) {
      s->idleCv.wait(lk, idle);
    } else if (!s->idleCv.wait_for(lk, std::chrono::duration<double>(timeout_s), idle)) {
      throw std::runtime_error("wait: timed out");
    }
    if (!s->workerError.empty()) throw std::runtime_error("server failed: " + s->workerError);
    return SUNDOG_OK;
  });
}

extern "C" int sundog_server_get_stats(sundog_server* s, sundog_server_stats* out) {
  return guardedS([&] {
    if (!s || !out) throw std::runtime_error("stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = sundog_server_stats{};
    out->paths = s->paths;
    if (s->rayCounter.p) {
      unsigned long long rays = 0;
      HIP_OK(hipMemcpy(&rays, s->rayCounter.p, 8, hipMemcpyDeviceToHost));
      out->rays_traced = rays;
    }
    out->batches = s->batches;
    out->work_items = s->items;
    out->kernel_ms = s->kernelMs;
    out->busy_ms = s->busyMs;
    out->used_bytes = s->used;
    out->budget_bytes = s->budget;
    out->active = (int)s->active.size();
    out->waiting = (int)s->waiting.size();
    out->tile_size = s->tileSize;
    out->waves = s->waves;
    out->pass_spp = s->passSpp;
    out->blocks_per_cu = s->blocksPerCU;
    out->stack = s->stackLds;
    out->graph_passes = s->graphPasses;
    out->graph_launches = s->graphLaunches;
    return SUNDOG_OK;
  });
}

// The per-request baseline: one request rendered alone on the resident
// scene with the megakernel, in chunks of 16 spp with a synchronisation
// after every chunk (the reference's launch pattern, src/capi_render.cpp:232-243).
extern "C" int sundog_server_render_serial(sundog_server* s, int scene_id, int width, int height, int spp,
                                           int64_t seed, float* beauty, double* ms_out) {
  return guardedS([&] {
    if (!s) throw std::runtime_error("serial: null handle");
    if (scene_id < 0 || scene_id >= (int)s->scenes.size()) throw std::runtime_error("serial: scene id out of range");
    ExclusiveUse hold(*s, "serial");
    const int serialWaves = hold.serialWaves;
    s->ensureDevice();
    ServerRequest r;
    r.scene = scene_id;
    r.width = width;
    r.height = height;
    r.spp = spp;
    r.seed = seed >= 0 ? (uint32_t)seed : s->scenes[scene_id]->P.seed;
    size_t n = (size_t)width * height * 16;
    r.accum.alloc(n);
    r.albedo.alloc(n);
    r.normal.alloc(n);
    HIP_OK(hipMemsetAsync(r.accum.p, 0, n, s->stream));
    HIP_OK(hipMemsetAsync(r.albedo.p, 0, n, s->stream));
    HIP_OK(hipMemsetAsync(r.normal.p, 0, n, s->stream));
    LaunchParams P = s->requestParams(r);
    sundog_session& sc = *s->scenes[scene_id];
    // spill room for the largest sample-group count a pass can use
    const size_t tiles = (size_t)((width + 15) / 16) * ((height + 15) / 16);
    size_t threads = tiles * kPoolSamples * renderBlockThreads();
    DevBuf sp;
    sp.alloc(std::max<size_t>(16, (size_t)sc.spillRows * threads * kStackEntryBytes));
    P.spill = sp.as<uint32_t>();
    P.spillStride = (int)threads;
    const size_t waves = tiles * (renderBlockThreads() / 64);
    DevBuf sb;
    sb.alloc(waves * (size_t)kPoolPlanes * kPoolSamples * 64 * sizeof(float));
    P.sampleBuf = sb.as<float>();
    P.sampleBufWaves = (int64_t)waves;
    HIP_OK(hipStreamSynchronize(s->stream));
    // serialWaves 0: the per-request choices a session makes (occupancy
    // variant, then sample groups, from timed passes)
    OccupancyPicker pick;
    GroupPicker gpick;
    auto launchTimed = [&](int offset, int count) {
      P.sampleOffset = offset;
      P.sppThisLaunch = count;
      bool trial = false, gTrial = false;
      const int w = serialWaves != 0 ? serialWaves : pick.pick(trial);
      int groups = 1;
      if (P.sppThisLaunch <= kPoolSamples) {
        if (trial || (serialWaves == 0 && pick.chosen == 0))
          groups = sampleGroupsFor(tiles, P.sppThisLaunch, s->numCUs);
        else
          groups = gpick.pick(gTrial, P.sppThisLaunch);
      }
      auto t0 = Clock::now();
      HIP_OK(launchRender(P, false, w, groups, s->stream));
      HIP_OK(hipStreamSynchronize(s->stream));
      const double msPerPath = std::chrono::duration<double, std::milli>(Clock::now() - t0).count() /
                               ((double)width * height * P.sppThisLaunch);
      if (trial) pick.record(w, msPerPath);
      if (gTrial) gpick.record(groups, msPerPath);
    };
    // 16-spp passes; while the occupancy variant is being timed, a large
    // frame's pass runs as quarter-pass trial launches, as a session's does
    auto t = Clock::now();
    for (int done = 0; done < spp; done += 16) {
      const int count = std::min(16, spp - done);
      const int sub = (count + 3) / 4;
      int d = 0;
      if (serialWaves == 0 && count >= 4 && (double)width * height * sub >= sundog_session::kTrialMinPaths)
        while (d < count && pick.chosen == 0) {
          const int c = std::min(sub, count - d);
          launchTimed(done + d, c);
          d += c;
        }
      if (d < count) launchTimed(done + d, count - d);
    }
    double ms = std::chrono::duration<double, std::milli>(Clock::now() - t).count();
    if (ms_out) *ms_out = ms;
    if (beauty) HIP_OK(hipMemcpy(beauty, r.accum.p, n, hipMemcpyDeviceToHost));
    return SUNDOG_OK;
  });
}

// Autotuner: coordinate descent over the persistent kernel's launch shape.
// Each trial renders the calibration request from sample 0 through the same
// batch path the worker uses and is timed with events on the server stream.
extern "C" int sundog_server_autotune(sundog_server* s, int scene_id, int width, int height, int spp,
                                      sundog_tune_result* out) {
  return guardedS([&] {
    if (!s) throw std::runtime_error("autotune: null handle");
    if (scene_id < 0 || scene_id >= (int)s->scenes.size()) throw std::runtime_error("autotune: scene id out of range");
    if (width <= 0 || height <= 0 || spp <= 0) throw std::runtime_error("autotune: width, height and spp must be positive");
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->started) throw std::runtime_error("autotune: call before the first submit");
    }
    ExclusiveUse hold(*s, "autotune");
    s->ensureDevice();
    ServerRequest r;
    r.scene = scene_id;
    r.width = width;
    r.height = height;
    r.spp = spp;
    r.seed = s->scenes[scene_id]->P.seed;
    size_t n = (size_t)width * height * 16;
    r.accum.alloc(n);
    r.albedo.alloc(n);
    r.normal.alloc(n);
    HIP_OK(hipMemsetAsync(r.accum.p, 0, n, s->stream));
    HIP_OK(hipMemsetAsync(r.albedo.p, 0, n, s->stream));
    HIP_OK(hipMemsetAsync(r.normal.p, 0, n, s->stream));
    const uint64_t paths0 = s->paths, batches0 = s->batches, items0 = s->items, graphs0 = s->graphLaunches;
    const double kms0 = s->kernelMs;
    int trials = 0;
    auto trial = [&]() {
      s->applyTuning();
      std::vector<ServerRequest*> batch{&r};
      r.done = 0;
      double ms = 0.0;
      while (r.done < r.spp) {
        double before = s->kernelMs;
        s->runBatch(batch, s->passesFor(batch));
        ms += s->kernelMs - before;
      }
      trials++;
      return (double)width * height * spp / (ms * 1e-3);
    };
    trial();  // warm-up: first launches of each kernel
    const double start = trial();
    double best = start;
    auto sweep = [&](int& knob, std::initializer_list<int> values) {
      int keep = knob;
      for (int v : values) {
        if (v == keep) continue;
        knob = v;
        double rate = trial();
        if (rate > best) {
          best = rate;
          keep = v;
        }
      }
      knob = keep;
    };
    sweep(s->waves, {2, 3, 4, 5});
    sweep(s->stackLds, {8, 12, 16, 24});
    sweep(s->tileSize, {8, 16, 32});
    const int fit = persistentBlocksPerCU(s->waves, s->stackLds);
    std::vector<int> depths;
    for (int b = fit; b >= 1; b--) depths.push_back(b);
    {
      int keep = s->blocksReq;
      for (int b : depths) {
        s->blocksReq = b;
        double rate = trial();
        if (rate > best) {
          best = rate;
          keep = b;
        }
      }
      s->blocksReq = keep;
    }
    sweep(s->graphPasses, {1, 2, 4, 8});
    s->applyTuning();
    // calibration work is not served work
    s->paths = paths0;
    s->batches = batches0;
    s->items = items0;
    s->graphLaunches = graphs0;
    s->kernelMs = kms0;
    if (out) {
      *out = sundog_tune_result{};
      out->tile = s->tileSize;
      out->waves = s->waves;
      out->stack = s->stackLds;
      out->blocks_per_cu = s->blocksPerCU;
      out->graph_passes = s->graphPasses;
      out->trials = trials;
      out->start_paths_per_s = start;
      out->best_paths_per_s = best;
    }
    return SUNDOG_OK;
  });
}
