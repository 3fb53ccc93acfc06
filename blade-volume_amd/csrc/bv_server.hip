// Serving loop: resident scenes, HBM-budget admission and continuous batching
// of view requests into persistent tile-queue launches.
//
// The reference's viewer renders one frame per loop turn, one dispatch per
// frame (blade-volume-view/src/bin/view.rs:593-840, RadFoamBackend::render,
// render.rs:413-456).  Here many clients submit views of resident scenes; a
// scheduler thread
//   1. admits pending requests in arrival order while a request slot is free
//      and the output image fits the HBM arena (admission by memory),
//   2. fills each launch round-robin with 8x8 ray tiles of every admitted
//      request (so requests that arrive while others run join the next launch
//      and their work overlaps on the device), one launch per kernel variant,
//   3. keeps up to two launches in flight on its own stream, and
//   4. when a request's last tile has run, copies its image to the client's
//      host buffer, frees the arena block and records the latency.
// Each ray's result depends only on the ray and the scene, so an image is the
// same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bv_host.h"

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation.
class Arena {
 public:
  void init(int64_t size) {
    size_ = size;
    free_.clear();
    free_[0] = size;
  }
  int64_t size() const { return size_; }
  bool alloc(int64_t bytes, int64_t* off) {
    bytes = (bytes + 255) & ~int64_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const int64_t rest = it->second - bytes, at = it->first + bytes;
        free_.erase(it);
        if (rest > 0) free_[at] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
      }
    }
    return false;
  }
  void release(int64_t off, int64_t bytes) {
    bytes = (bytes + 255) & ~int64_t(255);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t peak() const { return peak_; }

 private:
  int64_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<int64_t, int64_t> free_;
};

struct Request {
  uint64_t ticket = 0;
  bv_view_request req{};
  uint32_t tiles = 0, next_tile = 0, done_tiles = 0;
  int slot = -1;
  int64_t off = 0, bytes = 0;
  int variant = 0;
  Clock::time_point submitted;
  bool complete = false;
  double latency_ms = 0.0;
  hipEvent_t copied = nullptr;  // D2H copy of the image finished
};

struct Launch {
  hipEvent_t done = nullptr;
  std::vector<std::pair<uint64_t, uint32_t>> counts;  // ticket -> tiles in this launch
  std::vector<int> buffers;                           // tile buffers this launch holds until it is done
};

constexpr int kTileBuffers = 8;

}  // namespace

struct bv_server {
  bv_server_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  float* arena_mem = nullptr;
  Arena arena;
  bvk::ServeReq* d_reqs = nullptr;    // [max_active]
  bvk::ServeReq* h_reqs = nullptr;    // pinned staging of the table
  // Tile lists: pinned host staging + device copy + queue counter per buffer.
  // A buffer is reused only after the launch that used it has completed.
  uint32_t* d_tiles[kTileBuffers] = {};
  uint32_t* h_tiles[kTileBuffers] = {};
  uint32_t* d_counter[kTileBuffers] = {};
  std::vector<int> free_buffers;
  uint32_t tile_cap = 0;
  int cu_count = 256;
  std::vector<bv_scene*> scenes;
  std::vector<int> free_slots;

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> pending;
  std::vector<Request*> active;
  std::unordered_map<uint64_t, Request*> all;
  std::deque<Launch> inflight;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  int error = BV_OK;
  std::string error_msg;

  // statistics
  std::vector<double> latencies;
  uint64_t submitted = 0, completed = 0, rejected = 0, rays_done = 0;
  bool started = false;
  Clock::time_point first_submit, last_complete;
  double busy_ms = 0.0;
};

namespace {

int variant_of(const bv_server* sv, const bv_view_request& r) {
  const bv_scene* sc = sv->scenes[r.scene];
  return (r.renderer == BV_RENDER_SPLAT ? 16 : 0) + (r.mode == BV_EVAL_RELIGHT ? 8 : 0) + sc->sh_degree * 2 +
         (sc->bounded ? 1 : 0);
}

void fail(bv_server* sv, const std::string& msg) {
  sv->error = BV_ERR_HIP;
  sv->error_msg = msg;
}

// Called with the lock held: admit pending requests in arrival order.
void admit(bv_server* sv) {
  while (!sv->pending.empty() && !sv->free_slots.empty()) {
    Request* r = sv->pending.front();
    int64_t off;
    if (!sv->arena.alloc(r->bytes, &off)) break;  // wait for memory to free up
    sv->pending.pop_front();
    r->off = off;
    r->slot = sv->free_slots.back();
    sv->free_slots.pop_back();
    const bv_scene* sc = sv->scenes[r->req.scene];
    bvk::ServeReq& d = sv->h_reqs[r->slot];
    d.s = sc->view();
    d.bvh = r->req.renderer == BV_RENDER_SPLAT ? bvh::sphere_bvh_view(sc) : bvk::BvhView{};
    d.rs = bvh::camera_source(&r->req.camera, r->req.width, r->req.height, r->req.start_cell);
    d.st = bvh::settings_of(&r->req.settings);
    d.lit = r->req.lighting;
    d.out = sv->arena_mem + off / 4;
    hipError_t e = hipMemcpyAsync(sv->d_reqs + r->slot, &sv->h_reqs[r->slot], sizeof(bvk::ServeReq),
                                  hipMemcpyHostToDevice, sv->stream);
    if (e != hipSuccess) fail(sv, std::string("request table upload: ") + hipGetErrorString(e));
    sv->active.push_back(r);
  }
}

// Called with the lock held: build and enqueue one launch per variant.
void launch(bv_server* sv) {
  std::map<int, std::vector<Request*>> by_variant;
  for (Request* r : sv->active)
    if (r->next_tile < r->tiles) by_variant[r->variant].push_back(r);
  if (by_variant.empty() || sv->free_buffers.size() < by_variant.size()) return;
  Launch L;
  const uint32_t chunk = 64;  // tiles per request per round-robin turn
  for (auto& kv : by_variant) {
    const int b = sv->free_buffers.back();
    sv->free_buffers.pop_back();
    L.buffers.push_back(b);
    uint32_t* h = sv->h_tiles[b];
    uint32_t n = 0;
    bool more = true;
    while (more && n < sv->tile_cap) {
      more = false;
      for (Request* r : kv.second) {
        const uint32_t take = std::min({chunk, r->tiles - r->next_tile, sv->tile_cap - n});
        for (uint32_t k = 0; k < take; ++k) h[n++] = ((uint32_t)r->slot << 20) | (r->next_tile + k);
        if (take) L.counts.emplace_back(r->ticket, take);
        r->next_tile += take;
        if (r->next_tile < r->tiles) more = true;
        if (n >= sv->tile_cap) break;
      }
    }
    hipError_t e = hipMemcpyAsync(sv->d_tiles[b], h, n * sizeof(uint32_t), hipMemcpyHostToDevice, sv->stream);
    if (e != hipSuccess) fail(sv, std::string("tile upload: ") + hipGetErrorString(e));
    const bv_view_request& q = kv.second.front()->req;
    const bv_scene* sc = sv->scenes[q.scene];
    int rc;
    if (q.renderer == BV_RENDER_SPLAT) {
      // one-wave blocks; the LDS stack covers the deepest BVH among the variant's scenes
      uint32_t rows = 2;
      for (Request* r : kv.second) rows = std::max(rows, bvh::sphere_bvh_view(sv->scenes[r->req.scene]).stack_rows);
      const uint32_t blocks = std::min<uint32_t>(n, (uint32_t)sv->cu_count * 12);
      rc = bvh::launch_serve_splat(q.mode, sc->sh_degree, sv->d_reqs, sv->d_tiles[b], n, sv->d_counter[b], blocks,
                                   rows, sv->stream);
    } else {
      const uint32_t blocks = std::min<uint32_t>(n, (uint32_t)sv->cu_count * 8);
      rc = bvh::launch_serve(q.mode, sc->sh_degree, sc->bounded != 0, sv->d_reqs, sv->d_tiles[b], n,
                             sv->d_counter[b], blocks, sv->stream);
    }
    if (rc != BV_OK) fail(sv, "serve launch failed");
  }
  hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
  hipEventRecord(L.done, sv->stream);
  sv->inflight.push_back(std::move(L));
}

// Called with the lock held after the oldest launch finished.
void retire(bv_server* sv, Launch& L) {
  for (int b : L.buffers) sv->free_buffers.push_back(b);
  for (auto& tc : L.counts) {
    Request* r = sv->all[tc.first];
    r->done_tiles += tc.second;
    if (r->done_tiles == r->tiles) {
      const size_t bytes = (size_t)r->req.width * r->req.height * 16;
      hipError_t e = hipMemcpyAsync(r->req.host_rgba, sv->arena_mem + r->off / 4, bytes, hipMemcpyDeviceToHost,
                                    sv->stream);
      if (e != hipSuccess) fail(sv, std::string("result copy: ") + hipGetErrorString(e));
      hipEventCreateWithFlags(&r->copied, hipEventDisableTiming);
      hipEventRecord(r->copied, sv->stream);
    }
  }
}

// Called with the lock held: finish requests whose image has reached the host.
void finish_copies(bv_server* sv) {
  for (size_t i = 0; i < sv->active.size();) {
    Request* r = sv->active[i];
    if (r->copied && hipEventQuery(r->copied) == hipSuccess) {
      hipEventDestroy(r->copied);
      r->copied = nullptr;
      sv->arena.release(r->off, r->bytes);
      sv->free_slots.push_back(r->slot);
      r->complete = true;
      const auto now = Clock::now();
      r->latency_ms = std::chrono::duration<double, std::milli>(now - r->submitted).count();
      sv->latencies.push_back(r->latency_ms);
      sv->completed += 1;
      sv->rays_done += (uint64_t)r->req.width * r->req.height;
      sv->last_complete = now;
      sv->active.erase(sv->active.begin() + i);
    } else {
      ++i;
    }
  }
}

void worker_loop(bv_server* sv) {
  (void)hipSetDevice(sv->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    admit(sv);
    while (sv->inflight.size() < 2) {
      const size_t before = sv->inflight.size();
      launch(sv);
      if (sv->inflight.size() == before) break;
    }
    finish_copies(sv);
    sv->cv.notify_all();
    if (sv->stop && sv->pending.empty() && sv->active.empty() && sv->inflight.empty()) break;
    if (!sv->inflight.empty()) {
      hipEvent_t ev = sv->inflight.front().done;
      const auto t0 = Clock::now();
      lk.unlock();
      hipError_t e = hipEventSynchronize(ev);
      lk.lock();
      sv->busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
      if (e != hipSuccess) fail(sv, std::string("serve launch: ") + hipGetErrorString(e));
      Launch L = std::move(sv->inflight.front());
      sv->inflight.pop_front();
      retire(sv, L);
      hipEventDestroy(L.done);
    } else if (!sv->active.empty()) {
      // only image copies outstanding
      lk.unlock();
      (void)hipStreamSynchronize(sv->stream);
      lk.lock();
    } else {
      sv->cv.wait_for(lk, std::chrono::milliseconds(50));
    }
    if (sv->error != BV_OK) {
      // fail every waiting request rather than hang clients
      for (auto& kv : sv->all) kv.second->complete = true;
      sv->cv.notify_all();
      break;
    }
  }
}

}  // namespace

extern "C" {

int bv_server_create(const bv_server_config* cfg, bv_server** out) {
  BV_REQUIRE(cfg && out, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(cfg->hbm_budget_bytes >= (1 << 20), BV_ERR_INVALID, "budget below 1 MiB");
  BV_REQUIRE(cfg->max_active >= 1 && cfg->max_active <= 4095, BV_ERR_INVALID, "max_active must be 1..4095");
  *out = nullptr;
  bv_server* sv = new bv_server();
  sv->cfg = *cfg;
  sv->tile_cap = cfg->tiles_per_launch ? cfg->tiles_per_launch : 8192;
  BV_HIP(hipGetDevice(&sv->device));
  hipDeviceProp_t prop;
  BV_HIP(hipGetDeviceProperties(&prop, sv->device));
  sv->cu_count = prop.multiProcessorCount;
  BV_HIP(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
  hipError_t e = hipMalloc((void**)&sv->arena_mem, (size_t)cfg->hbm_budget_bytes);
  if (e != hipSuccess) {
    bvh::set_error("server arena allocation failed");
    delete sv;
    return BV_ERR_OOM;
  }
  sv->arena.init(cfg->hbm_budget_bytes);
  BV_HIP(hipMalloc((void**)&sv->d_reqs, sizeof(bvk::ServeReq) * cfg->max_active));
  BV_HIP(hipHostMalloc((void**)&sv->h_reqs, sizeof(bvk::ServeReq) * cfg->max_active, hipHostMallocDefault));
  for (int b = 0; b < kTileBuffers; ++b) {
    BV_HIP(hipMalloc((void**)&sv->d_tiles[b], sizeof(uint32_t) * sv->tile_cap));
    BV_HIP(hipHostMalloc((void**)&sv->h_tiles[b], sizeof(uint32_t) * sv->tile_cap, hipHostMallocDefault));
    BV_HIP(hipMalloc((void**)&sv->d_counter[b], sizeof(uint32_t)));
    sv->free_buffers.push_back(b);
  }
  for (int s = (int)cfg->max_active - 1; s >= 0; --s) sv->free_slots.push_back(s);
  sv->worker = std::thread(worker_loop, sv);
  *out = sv;
  return BV_OK;
}

int bv_server_add_scene(bv_server* sv, bv_scene* scene, uint32_t* id) {
  BV_REQUIRE(sv && scene && id, BV_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->scenes.push_back(scene);
  *id = (uint32_t)(sv->scenes.size() - 1);
  return BV_OK;
}

int bv_server_submit(bv_server* sv, const bv_view_request* req, uint64_t* ticket) {
  BV_REQUIRE(sv && req && ticket, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(req->width > 0 && req->height > 0 && req->width <= 16384 && req->height <= 16384, BV_ERR_INVALID,
             "image size out of range");
  BV_REQUIRE(req->host_rgba != nullptr, BV_ERR_INVALID, "null result buffer");
  BV_REQUIRE((uint64_t)((req->width + 7) / 8) * ((req->height + 7) / 8) <= (1u << 20), BV_ERR_INVALID,
             "image has more than 2^20 8x8 tiles");
  BV_REQUIRE(req->mode == BV_EVAL_SH || req->mode == BV_EVAL_RELIGHT, BV_ERR_INVALID, "mode must be SH or RELIGHT");
  BV_REQUIRE(req->renderer == BV_RENDER_WALK || req->renderer == BV_RENDER_SPLAT, BV_ERR_INVALID,
             "renderer must be BV_RENDER_WALK or BV_RENDER_SPLAT");
  std::unique_lock<std::mutex> lk(sv->mu);
  if (sv->error != BV_OK) {
    bvh::set_error(sv->error_msg);
    return sv->error;
  }
  BV_REQUIRE(req->scene < sv->scenes.size(), BV_ERR_INVALID, "unknown scene id");
  const bv_scene* sc = sv->scenes[req->scene];
  if (req->mode == BV_EVAL_RELIGHT) {
    BV_REQUIRE(sc->material != nullptr, BV_ERR_INVALID, "relighting needs a scene with materials");
    BV_REQUIRE(req->lighting.count >= 1 && req->lighting.count <= BV_MAX_LIGHTS, BV_ERR_INVALID,
               "lighting needs 1..8 lights");
  }
  bv_view_request rq = *req;
  if (rq.renderer == BV_RENDER_SPLAT) {
    BV_REQUIRE(sc->bounded, BV_ERR_UNSUPPORTED, "PowerFoam splats require radii");
    rq.start_cell = 0;
    if (!sc->bvh_valid) {
      lk.unlock();
      const int rc = bvh::ensure_sphere_bvh(const_cast<bv_scene*>(sc), nullptr);
      if (rc != BV_OK) return rc;
      lk.lock();
    }
  } else if (rq.start_cell == BV_NO_CELL) {
    lk.unlock();
    const int rc = bv_containing_cell(const_cast<bv_scene*>(sc), rq.camera.position, &rq.start_cell, nullptr);
    if (rc != BV_OK) return rc;
    lk.lock();
  }
  BV_REQUIRE(rq.start_cell < sc->n, BV_ERR_INVALID, "start_cell out of bounds");
  const int64_t bytes = (int64_t)req->width * req->height * 16;
  if (bytes > sv->arena.size()) {
    sv->rejected += 1;
    bvh::set_error("request image exceeds the server's HBM budget");
    return BV_ERR_CAPACITY;
  }
  Request* r = new Request();
  r->ticket = sv->next_ticket++;
  r->req = rq;
  r->bytes = bytes;
  r->tiles = ((req->width + 7) / 8) * ((req->height + 7) / 8);
  r->variant = variant_of(sv, *req);
  r->submitted = Clock::now();
  if (!sv->started) {
    sv->started = true;
    sv->first_submit = r->submitted;
  }
  sv->submitted += 1;
  sv->all[r->ticket] = r;
  sv->pending.push_back(r);
  *ticket = r->ticket;
  lk.unlock();
  sv->cv.notify_all();
  return BV_OK;
}

int bv_server_poll(bv_server* sv, uint64_t ticket, int* done) {
  BV_REQUIRE(sv && done, BV_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->all.find(ticket);
  BV_REQUIRE(it != sv->all.end(), BV_ERR_INVALID, "unknown ticket");
  *done = it->second->complete ? 1 : 0;
  return sv->error;
}

int bv_server_wait(bv_server* sv, uint64_t ticket, double* latency_ms) {
  BV_REQUIRE(sv != nullptr, BV_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->all.find(ticket);
  BV_REQUIRE(it != sv->all.end(), BV_ERR_INVALID, "unknown ticket");
  Request* r = it->second;
  sv->cv.wait(lk, [&] { return r->complete || sv->error != BV_OK; });
  if (sv->error != BV_OK) {
    bvh::set_error(sv->error_msg);
    return sv->error;
  }
  if (latency_ms) *latency_ms = r->latency_ms;
  return BV_OK;
}

int bv_server_get_stats(bv_server* sv, bv_server_stats* out) {
  BV_REQUIRE(sv && out, BV_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = sv->submitted;
  out->completed = sv->completed;
  out->rejected = sv->rejected;
  out->peak_bytes = sv->arena.peak();
  if (!sv->latencies.empty()) {
    std::vector<double> l = sv->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) {
      const double x = p * (double)(l.size() - 1);
      const size_t i = (size_t)x;
      const double f = x - (double)i;
      return i + 1 < l.size() ? l[i] * (1.0 - f) + l[i + 1] * f : l[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->max_ms = l.back();
    double s = 0.0;
    for (double v : l) s += v;
    out->mean_ms = s / (double)l.size();
    const double wall = std::chrono::duration<double>(sv->last_complete - sv->first_submit).count();
    out->mrays_per_s = wall > 0 ? (double)sv->rays_done / wall / 1e6 : 0.0;
    out->busy_ms = sv->busy_ms;
  }
  return BV_OK;
}

int bv_server_destroy(bv_server* sv) {
  if (!sv) return BV_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipStreamSynchronize(sv->stream);
  for (auto& kv : sv->all) {
    if (kv.second->copied) (void)hipEventDestroy(kv.second->copied);
    delete kv.second;
  }
  for (int b = 0; b < kTileBuffers; ++b) {
    if (sv->d_tiles[b]) (void)hipFree(sv->d_tiles[b]);
    if (sv->h_tiles[b]) (void)hipHostFree(sv->h_tiles[b]);
    if (sv->d_counter[b]) (void)hipFree(sv->d_counter[b]);
  }
  if (sv->d_reqs) (void)hipFree(sv->d_reqs);
  if (sv->h_reqs) (void)hipHostFree(sv->h_reqs);
  if (sv->arena_mem) (void)hipFree(sv->arena_mem);
  if (sv->stream) (void)hipStreamDestroy(sv->stream);
  delete sv;
  return BV_OK;
}

}  // extern "C"
