// server.hip - the serving form: resident scenes, continuous batching, in-flight slots.
//
// One worker thread owns all device work.  It loops:
//   1. for every busy slot whose completion event has fired: mark its requests done (or, if the
//      batch overflowed the slot's pair capacity, grow it and run the same batch again);
//   2. while a slot is free and requests are queued: take queued requests in arrival order into a
//      batch (up to max_views_per_batch frames and max_splats_per_batch splats), admitting their
//      scenes (upload, evicting least recently used scenes that no queued batch uses, within the
//      byte budget), and enqueue the batch on the slot's stream through Engine::run_graph;
//   3. otherwise sleep until a submit or a short timeout.
// Requests that arrive while batches run are therefore batched as soon as a slot frees, and the
// slots' streams overlap on the device.  A batch is one Engine launch sequence over all its
// frames, so a request's output depends only on its own scene and camera.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "hdsr.h"

using namespace hdsr;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint32_t sh_stride_of(uint32_t deg) { return sh_row_floats(deg); }

struct Scene {
  hdsr_splats host{};
  uint32_t n = 0, sh_degree = 0;
  bool resident = false;
  // device copy: packed records (the projection's only splat input, built at upload), the
  // normals and the order
  float4* packed = nullptr;
  float* normals = nullptr;
  uint32_t* sorted = nullptr;
  uint64_t bytes = 0;
  uint64_t last_use = 0;
  uint32_t inflight = 0;  // requests of this scene that are batched and not yet complete
};

struct OutBuf {
  float* rgba = nullptr;
  float* depth = nullptr;
};

struct Req {
  uint32_t scene = 0, w = 0, h = 0;
  uint8_t uni[HDSR_FRAME_UNIFORM_BYTES];
  double t_submit = 0, t_done = 0;
  uint32_t batch_views = 0;
  bool done = false;
  OutBuf out;
};

struct Slot {
  Engine eng;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<uint64_t> tickets;
  std::vector<FrameDesc> frames;
};

}  // namespace

struct hdsr_server {
  hdsr_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Req> reqs;
  std::unordered_map<uint32_t, Scene> scenes;
  std::vector<std::unique_ptr<Slot>> slots;
  std::multimap<std::pair<uint32_t, uint32_t>, OutBuf> pool;
  uint64_t next_ticket = 1, use_clock = 0, resident_bytes = 0;
  uint64_t pending = 0;  // submitted and not complete
  // statistics since the last reset
  uint64_t frames = 0, batches = 0, graph_replays = 0, uploads = 0, evictions = 0, batch_views_sum = 0;
  std::vector<double> latencies;
  double t_first = 0, t_last = 0;
  std::thread worker;
  bool stop = false;
  bool held = false;  // hdsr_server_hold: no new batch is formed (batches in flight complete)
  std::string err;

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }

  void free_scene_device(Scene& s) {
    void* p[] = {s.normals, s.sorted, s.packed};
    for (void* q : p)
      if (q) (void)hipFree(q);
    s.normals = nullptr;
    s.sorted = nullptr;
    s.packed = nullptr;
    if (s.resident) resident_bytes -= s.bytes;
    s.resident = false;
  }

  template <typename T>
  bool up(T** dst, const T* src, size_t count) {
    *dst = nullptr;
    if (!count) return true;
    if (hipMalloc((void**)dst, count * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
  }

  // Makes s resident, evicting least recently used idle scenes if the budget requires it.
  bool admit(Scene& s) {
    if (s.resident) return true;
    const uint64_t budget = cfg.hbm_budget_bytes;
    while (budget && resident_bytes + s.bytes > budget) {
      Scene* lru = nullptr;
      for (auto& kv : scenes) {
        Scene& c = kv.second;
        if (c.resident && c.inflight == 0 && (!lru || c.last_use < lru->last_use)) lru = &c;
      }
      if (!lru) return false;
      free_scene_device(*lru);
      ++evictions;
    }
    const size_t n = s.n;
    const uint32_t stride = sh_stride_of(s.sh_degree);
    // attribute arrays go up into temporaries and are packed on the device
    float *pos = nullptr, *scale = nullptr, *rot = nullptr, *opac = nullptr, *sh = nullptr;
    uint32_t* material = nullptr;
    bool ok = up(&pos, s.host.positions, n * 3) && up(&scale, s.host.scales, n * 3) &&
              up(&rot, s.host.rotations, n * 4) && up(&opac, s.host.opacities, n) &&
              up(&material, s.host.materials, n) && up(&sh, s.host.sh_data, n * stride) &&
              up(&s.normals, s.host.normals, (size_t)s.host.normals_len);
    if (ok && n) {
      ok = hipMalloc((void**)&s.packed, n * (size_t)packed_stride4(stride) * sizeof(float4)) == hipSuccess;
      if (ok) {
        launch_pack_scene(pos, scale, rot, opac, material, sh, (uint32_t)n, stride, s.packed, nullptr);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(nullptr) == hipSuccess;
      } else {
        s.packed = nullptr;
        (void)hipGetLastError();
      }
    }
    void* tmp[] = {pos, scale, rot, opac, sh, material};
    for (void* q : tmp)
      if (q) (void)hipFree(q);
    if (ok) {
      if (s.host.sorted_indices) {
        ok = up(&s.sorted, s.host.sorted_indices, n);
      } else {
        std::vector<uint32_t> id(n);
        for (size_t i = 0; i < n; ++i) id[i] = (uint32_t)i;
        ok = up(&s.sorted, id.data(), n);
      }
    }
    if (!ok) {
      s.resident = true;  // so free_scene_device balances resident_bytes
      resident_bytes += s.bytes;
      free_scene_device(s);
      return false;
    }
    s.resident = true;
    resident_bytes += s.bytes;
    ++uploads;
    return true;
  }

  OutBuf take_out(uint32_t w, uint32_t h) {
    auto it = pool.find({w, h});
    if (it != pool.end()) {
      OutBuf o = it->second;
      pool.erase(it);
      return o;
    }
    OutBuf o;
    const size_t px = (size_t)w * h;
    if (hipMalloc((void**)&o.rgba, px * 16) != hipSuccess || hipMalloc((void**)&o.depth, px * 4) != hipSuccess) {
      (void)hipGetLastError();
      if (o.rgba) (void)hipFree(o.rgba);
      return OutBuf{};
    }
    return o;
  }

  // Forms a batch from the head of the queue and enqueues it on 'slot'.  Called with mu held.
  void launch(Slot& slot) {
    const uint32_t maxv = cfg.max_views_per_batch;
    const uint64_t maxs = cfg.max_splats_per_batch;
    slot.tickets.clear();
    slot.frames.clear();
    uint64_t splats = 0;
    while (!queue.empty() && slot.tickets.size() < maxv) {
      const uint64_t t = queue.front();
      Req& r = reqs[t];
      auto sit = scenes.find(r.scene);
      if (sit == scenes.end()) {  // scene removed after submit: complete with no output
        queue.pop_front();
        r.done = true;
        r.t_done = now_s();
        --pending;
        cv_done.notify_all();
        continue;
      }
      Scene& s = sit->second;
      if (!slot.tickets.empty() && splats + s.n > maxs) break;
      s.inflight++;  // pins the scene against eviction while admitting the rest of the batch
      if (!admit(s)) {
        s.inflight--;
        break;
      }
      if (!r.out.rgba) r.out = take_out(r.w, r.h);
      if (!r.out.rgba) {
        s.inflight--;
        break;
      }
      s.last_use = ++use_clock;
      FrameDesc F{};
      parse_uniforms(r.uni, &F);
      F.sorted = s.sorted;
      F.normals = s.normals;
      F.normals_len = (uint32_t)std::min<uint64_t>(s.host.normals_len, 0xFFFFFFFFull);
      F.n = s.n;
      F.sh_degree = s.sh_degree;
      F.sh_stride = sh_stride_of(s.sh_degree);
      F.packed = s.packed;
      F.pk_stride4 = packed_stride4(F.sh_stride);
      F.width = r.w;
      F.height = r.h;
      F.rgba = r.out.rgba;
      F.out_depth = r.out.depth;
      slot.frames.push_back(F);
      slot.tickets.push_back(t);
      splats += s.n;
      queue.pop_front();
    }
    if (slot.tickets.empty()) return;
    enqueue(slot);
  }

  void enqueue(Slot& slot) {
    const uint32_t V = (uint32_t)slot.frames.size();
    int rc;
    if (cfg.use_graphs) {
      bool replayed = false;
      rc = slot.eng.run_graph(slot.frames.data(), V, slot.stream, false, &replayed);
      if (replayed) ++graph_replays;
    } else {
      rc = slot.eng.run(slot.frames.data(), V, slot.stream, false, false);
    }
    if (rc == 0 && hipEventRecord(slot.done, slot.stream) != hipSuccess) rc = HDSR_E_HIP;
    if (rc) {
      err = slot.eng.err;
      // fail the batch: complete its requests without output
      for (uint64_t t : slot.tickets) finish(t, 0);
      slot.tickets.clear();
      return;
    }
    ++batches;
    batch_views_sum += V;
    slot.busy = true;
  }

  void finish(uint64_t t, uint32_t views) {
    Req& r = reqs[t];
    r.done = true;
    r.t_done = now_s();
    r.batch_views = views;
    auto sit = scenes.find(r.scene);
    if (sit != scenes.end() && sit->second.inflight) sit->second.inflight--;
    latencies.push_back((r.t_done - r.t_submit) * 1e3);
    ++frames;
    t_last = r.t_done;
    --pending;
  }

  void complete(Slot& slot) {
    if (slot.eng.overflowed()) {
      slot.eng.grow_for_overflow();
      enqueue(slot);  // the same frames again, with room for all pairs
      return;
    }
    slot.busy = false;
    const uint32_t V = (uint32_t)slot.tickets.size();
    for (uint64_t t : slot.tickets) finish(t, V);
    slot.tickets.clear();
    cv_done.notify_all();
  }

  void loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      bool any_busy = false;
      for (auto& sp : slots) {
        Slot& s = *sp;
        if (s.busy && hipEventQuery(s.done) == hipSuccess) complete(s);
        any_busy |= s.busy;
      }
      bool launched = false;
      for (auto& sp : slots) {
        if (queue.empty() || (held && !stop)) break;
        if (!sp->busy) {
          launch(*sp);
          launched |= sp->busy;
        }
      }
      if (stop && queue.empty() && !any_busy && !launched) break;
      if (launched) continue;
      if (any_busy)
        cv_work.wait_for(lk, std::chrono::microseconds(50));
      else
        cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
};

extern "C" {

int hdsr_server_create(const hdsr_server_config* cfg, hdsr_server** out) {
  if (!out) return HDSR_E_INVALID_ARG;
  *out = nullptr;
  std::unique_ptr<hdsr_server> s(new hdsr_server());
  if (cfg) s->cfg = *cfg;
  if (!s->cfg.max_views_per_batch || s->cfg.max_views_per_batch > kMaxFrames) s->cfg.max_views_per_batch = kMaxFrames;
  if (!s->cfg.max_splats_per_batch) s->cfg.max_splats_per_batch = 32ull << 20;
  if (!s->cfg.slots) s->cfg.slots = 2;
  if (s->cfg.slots > 8) s->cfg.slots = 8;
  if (hipGetDevice(&s->device) != hipSuccess) return HDSR_E_HIP;
  for (uint32_t i = 0; i < s->cfg.slots; ++i) {
    std::unique_ptr<Slot> sl(new Slot());
    if (hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl->done, hipEventDisableTiming) != hipSuccess)
      return HDSR_E_HIP;
    sl->eng.set_min_capacity(64ull << 20);
    s->slots.push_back(std::move(sl));
  }
  hdsr_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->loop(); });
  *out = raw;
  return HDSR_OK;
}

void hdsr_server_destroy(hdsr_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipDeviceSynchronize();
  for (auto& kv : s->scenes) s->free_scene_device(kv.second);
  for (auto& kv : s->reqs) {
    if (kv.second.out.rgba) (void)hipFree(kv.second.out.rgba);
    if (kv.second.out.depth) (void)hipFree(kv.second.out.depth);
  }
  for (auto& kv : s->pool) {
    (void)hipFree(kv.second.rgba);
    (void)hipFree(kv.second.depth);
  }
  for (auto& sl : s->slots) {
    (void)hipStreamDestroy(sl->stream);
    (void)hipEventDestroy(sl->done);
  }
  delete s;
}

int hdsr_server_add_scene(hdsr_server* s, uint32_t id, const hdsr_splats* sp) {
  if (!s || !sp) return HDSR_E_INVALID_ARG;
  if (sp->memory != HDSR_MEM_HOST) return s->fail(HDSR_E_INVALID_ARG, "server scenes are registered from host memory");
  uint32_t deg = 0;
  const char* why = "";
  if (!resolve_sh_degree(sp->count, sp->sh_degree, sp->sh_data_len, &deg, &why))
    return s->fail(HDSR_E_INVALID_ARG, why);
  if (sp->count && (!sp->positions || !sp->scales || !sp->rotations || !sp->opacities || !sp->materials ||
                    !sp->sh_data))
    return s->fail(HDSR_E_INVALID_ARG, "missing splat array");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->scenes.find(id);
  if (it != s->scenes.end()) {
    if (it->second.inflight) return s->fail(HDSR_E_STATE, "scene is in use by queued requests");
    s->free_scene_device(it->second);
    s->scenes.erase(it);
  }
  Scene sc;
  sc.host = *sp;
  sc.host.sh_degree = deg;
  sc.n = sp->count;
  sc.sh_degree = deg;
  sc.bytes = (uint64_t)sp->count * (4 + 16ull * packed_stride4(sh_stride_of(deg))) + sp->normals_len * 4;
  if (s->cfg.hbm_budget_bytes && sc.bytes > s->cfg.hbm_budget_bytes)
    return s->fail(HDSR_E_CAPACITY, "scene larger than the whole budget");
  s->scenes[id] = sc;
  return HDSR_OK;
}

int hdsr_server_remove_scene(hdsr_server* s, uint32_t id) {
  if (!s) return HDSR_E_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->scenes.find(id);
  if (it == s->scenes.end()) return s->fail(HDSR_E_INVALID_ARG, "unknown scene");
  if (it->second.inflight) return s->fail(HDSR_E_STATE, "scene is in use by queued requests");
  s->free_scene_device(it->second);
  s->scenes.erase(it);
  return HDSR_OK;
}

int hdsr_server_submit(hdsr_server* s, const hdsr_request* q, uint64_t* ticket) {
  if (!s || !q || !q->uniforms256 || !ticket) return HDSR_E_INVALID_ARG;
  const uint32_t* u = static_cast<const uint32_t*>(q->uniforms256);
  if (q->width == 0 || q->height == 0 || u[20] != (q->width + 15) / 16 || u[21] != (q->height + 15) / 16 ||
      u[20] > 255 || u[21] > 255)
    return s->fail(HDSR_E_INVALID_ARG, "uniform tile grid does not match the request size");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->scenes.find(q->scene_id);
  if (it == s->scenes.end()) return s->fail(HDSR_E_INVALID_ARG, "unknown scene");
  if (u[22] != it->second.n || u[25] != it->second.sh_degree)
    return s->fail(HDSR_E_INVALID_ARG, "uniform splatCount / shDegree differ from the scene");
  const uint64_t t = s->next_ticket++;
  Req& r = s->reqs[t];
  r.scene = q->scene_id;
  r.w = q->width;
  r.h = q->height;
  memcpy(r.uni, q->uniforms256, HDSR_FRAME_UNIFORM_BYTES);
  r.t_submit = now_s();
  if (s->pending == 0 && s->frames == 0) s->t_first = r.t_submit;
  ++s->pending;
  s->queue.push_back(t);
  *ticket = t;
  s->cv_work.notify_one();
  return HDSR_OK;
}

int hdsr_server_hold(hdsr_server* s, uint32_t hold) {
  if (!s) return HDSR_E_INVALID_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->held = hold != 0;
  }
  s->cv_work.notify_all();
  return HDSR_OK;
}

int hdsr_server_wait(hdsr_server* s, uint64_t t, double timeout_s) {
  if (!s) return HDSR_E_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->held) {  // a held server would never complete a queued ticket
    s->held = false;
    s->cv_work.notify_all();
  }
  auto pred = [&] {
    auto it = s->reqs.find(t);
    return it == s->reqs.end() || it->second.done;
  };
  if (timeout_s > 0) {
    if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
      return s->fail(HDSR_E_STATE, "timed out");
  } else {
    s->cv_done.wait(lk, pred);
  }
  return s->reqs.count(t) ? HDSR_OK : s->fail(HDSR_E_INVALID_ARG, "unknown ticket");
}

int hdsr_server_result(hdsr_server* s, uint64_t t, hdsr_result_info* info) {
  if (!s || !info) return HDSR_E_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(t);
  if (it == s->reqs.end()) return s->fail(HDSR_E_INVALID_ARG, "unknown ticket");
  const Req& r = it->second;
  info->done = r.done;
  info->width = r.w;
  info->height = r.h;
  info->batch_views = r.batch_views;
  info->latency_ms = r.done ? (r.t_done - r.t_submit) * 1e3 : 0.0;
  info->rgba = r.out.rgba;
  info->depth = r.out.depth;
  return HDSR_OK;
}

int hdsr_server_read(hdsr_server* s, uint64_t t, int which, void* host, uint64_t bytes) {
  if (!s || !host) return HDSR_E_INVALID_ARG;
  float* src = nullptr;
  size_t need = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(t);
    if (it == s->reqs.end()) return s->fail(HDSR_E_INVALID_ARG, "unknown ticket");
    const Req& r = it->second;
    if (!r.done || !r.out.rgba) return s->fail(HDSR_E_STATE, "result not ready");
    const size_t px = (size_t)r.w * r.h;
    if (which == HDSR_OUT_RGBA) {
      src = r.out.rgba;
      need = px * 16;
    } else if (which == HDSR_OUT_DEPTH) {
      src = r.out.depth;
      need = px * 4;
    } else {
      return s->fail(HDSR_E_INVALID_ARG, "server results hold RGBA and depth");
    }
  }
  if (bytes < need) return s->fail(HDSR_E_INVALID_ARG, "host buffer too small");
  if (hipMemcpy(host, src, need, hipMemcpyDeviceToHost) != hipSuccess) return s->fail(HDSR_E_HIP, "readback failed");
  return HDSR_OK;
}

int hdsr_server_release(hdsr_server* s, uint64_t t) {
  if (!s) return HDSR_E_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(t);
  if (it == s->reqs.end()) return s->fail(HDSR_E_INVALID_ARG, "unknown ticket");
  if (!it->second.done) return s->fail(HDSR_E_STATE, "request still in flight");
  if (it->second.out.rgba) s->pool.insert({{it->second.w, it->second.h}, it->second.out});
  s->reqs.erase(it);
  return HDSR_OK;
}

int hdsr_server_drain(hdsr_server* s) {
  if (!s) return HDSR_E_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->held) {
    s->held = false;
    s->cv_work.notify_all();
  }
  s->cv_done.wait(lk, [&] { return s->pending == 0; });
  return HDSR_OK;
}

int hdsr_server_get_stats(hdsr_server* s, hdsr_server_stats* o) {
  if (!s || !o) return HDSR_E_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  memset(o, 0, sizeof(*o));
  o->frames = s->frames;
  o->batches = s->batches;
  o->graph_replays = s->graph_replays;
  o->uploads = s->uploads;
  o->evictions = s->evictions;
  o->resident_bytes = s->resident_bytes;
  for (auto& kv : s->scenes) o->resident_scenes += kv.second.resident;
  o->queued = (uint32_t)s->queue.size();
  if (!s->latencies.empty()) {
    std::vector<double> v = s->latencies;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5))]; };
    o->p50_ms = pct(0.50);
    o->p99_ms = pct(0.99);
  }
  o->mean_batch_views = s->batches ? (double)s->batch_views_sum / s->batches : 0.0;
  o->busy_s = s->frames ? s->t_last - s->t_first : 0.0;
  return HDSR_OK;
}

int hdsr_server_reset_stats(hdsr_server* s) {
  if (!s) return HDSR_E_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->frames = s->batches = s->graph_replays = s->uploads = s->evictions = s->batch_views_sum = 0;
  s->latencies.clear();
  s->t_first = s->t_last = 0;
  return HDSR_OK;
}

const char* hdsr_server_last_error(hdsr_server* s) { return s ? s->err.c_str() : "null server"; }

}  // extern "C"
