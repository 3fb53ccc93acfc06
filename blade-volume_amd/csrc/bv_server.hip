// Serving loop: resident scenes, memory-budget admission and continuous
// batching of view requests into one persistent ray-tile kernel.
//
// The reference's viewer renders one frame per loop turn, one dispatch per
// frame (blade-volume-view/src/bin/view.rs:593-840, RadFoamBackend::render,
// render.rs:413-456).  Here many clients submit views of resident scenes, and
// one persistent kernel (bvk::ServeQueue) serves them all: its blocks take 8x8
// ray tiles from a ring in pinned host memory that a scheduler thread keeps
// filling while the kernel runs, so a request that arrives while others run
// has its tiles in the ring within one scheduler turn (about 20 us) and they
// run among the tiles already there.  The scheduler thread
//   1. admits pending requests in arrival order while a request slot is free
//      and the output image fits the result arena (admission by memory), and
//      writes each admitted request's descriptor to its slot;
//   2. keeps up to `tiles_ahead` published tiles unfinished, taking up to 64
//      tiles from each admitted request in turn, fewest remaining first, so
//      small requests finish among large ones;
//   3. reads the per-entry completion words the kernel writes, and hands a
//      request whose last tile is done to the copier thread, which copies the
//      image from the result arena (pinned host memory the kernel writes
//      directly) to the client's buffer and records the latency;
//   4. runs one kernel variant (renderer, mode, SH degree, boundedness) at a
//      time: when the running variant has no work left and another waits, or
//      after 2 ms without work, it stops the kernel (a flag in the ring's
//      control block) and, when needed, starts the next variant's.
// The kernel's grid leaves one block slot per CU free, so other device work
// (the containing-cell query of a walk request, other streams' launches) still
// finds room while it runs.  Each ray's result depends only on the ray and
// the scene, so an image is the same alone and in any mix.
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

// First-fit allocator over one allocation.
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
};

constexpr uint32_t kRingEntries = 1u << 15;       // ring of tile codes (a power of two)
constexpr int kIdleStopUs = 2000;                 // stop the kernel after this long without work
constexpr int kTurnUs = 20;                       // scheduler turn while tiles are outstanding
constexpr uint32_t kQuantum = 64;                 // tiles of one request per round-robin turn
constexpr double kHeartbeatTimeoutS = 2.0;        // the kernel leaves if the scheduler stalls this long

__global__ void k_serve_reset(uint64_t* claimed, uint64_t* published_dev, uint32_t* stop_dev, uint64_t published) {
  *claimed = published;
  *published_dev = published;
  *stop_dev = 0u;
}

}  // namespace

struct bv_server {
  bv_server_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 256;
  uint64_t dead_ticks = 0;
  // pinned host memory shared with the kernel (host pointers; device_* are
  // the same memory as the kernel addresses it)
  void* pinned = nullptr;
  bvk::ServeCtl* ctl = nullptr;
  uint32_t* ring = nullptr;
  uint32_t* done = nullptr;
  bvk::ServeReq* reqs = nullptr;
  char* results = nullptr;
  char* device_base = nullptr;
  size_t off_ring = 0, off_done = 0, off_reqs = 0, off_results = 0;
  Arena arena;
  // device counters of the running kernel
  uint64_t* d_counters = nullptr;  // [0] claimed, [1] published mirror; then the stop mirror
  // ring bookkeeping (host side): entries published, the oldest one not yet
  // done, and how many published entries are not done
  uint64_t published = 0, frontier = 0, outstanding = 0;
  std::vector<Request*> entry_req;
  std::vector<uint8_t> entry_done;
  // the running kernel
  bool running = false;
  int run_variant = -1;
  uint32_t run_rows = 0;  // the running splat kernel's LDS stack rows
  uint32_t tiles_ahead = 0;
  hipEvent_t kernel_done = nullptr;
  Clock::time_point idle_since;
  bool idle = true;
  std::vector<bv_scene*> scenes;
  std::vector<int> free_slots;

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> pending;
  std::vector<Request*> active;
  std::unordered_map<uint64_t, Request*> all;
  std::deque<Request*> to_copy;
  std::condition_variable copy_cv;
  uint64_t next_ticket = 1;
  uint32_t beat = 0;
  bool stop = false;
  bool copier_stop = false;
  std::thread worker, copier;
  int error = BV_OK;
  std::string error_msg;

  // statistics
  std::vector<double> latencies;
  uint64_t submitted = 0, completed = 0, rejected = 0, rays_done = 0;
  bool started = false;
  Clock::time_point first_submit, last_complete;
  double busy_ms = 0.0;
  Clock::time_point run_started;
};

namespace {

// variant code: renderer, mode, SH degree, boundedness
int variant_of(const bv_server* sv, const bv_view_request& r) {
  const bv_scene* sc = sv->scenes[r.scene];
  return (r.renderer == BV_RENDER_SPLAT ? 16 : 0) + (r.mode == BV_EVAL_RELIGHT ? 8 : 0) + sc->sh_degree * 2 +
         (sc->bounded ? 1 : 0);
}

void fail(bv_server* sv, const std::string& msg) {
  sv->error = BV_ERR_HIP;
  sv->error_msg = msg;
}

template <class T>
T* dev_ptr(const bv_server* sv, size_t off) {
  return reinterpret_cast<T*>(sv->device_base + off);
}

bvk::ServeQueue queue_of(const bv_server* sv) {
  bvk::ServeQueue q;
  q.ctl = dev_ptr<const bvk::ServeCtl>(sv, 0);
  q.ring = dev_ptr<const uint32_t>(sv, sv->off_ring);
  q.reqs = dev_ptr<const bvk::ServeReq>(sv, sv->off_reqs);
  q.done = dev_ptr<uint32_t>(sv, sv->off_done);
  q.claimed = sv->d_counters;
  q.published_dev = sv->d_counters + 1;
  q.stop_dev = reinterpret_cast<uint32_t*>(sv->d_counters + 2);
  q.mask = kRingEntries - 1;
  q.dead_ticks = sv->dead_ticks;
  return q;
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
    bvk::ServeReq d;
    d.s = sc->view();
    d.bvh = r->req.renderer == BV_RENDER_SPLAT ? bvh::sphere_bvh_view(sc) : bvk::BvhView{};
    d.rs = bvh::camera_source(&r->req.camera, r->req.width, r->req.height, r->req.start_cell);
    d.st = bvh::settings_of(&r->req.settings);
    d.lit = r->req.lighting;
    d.out = dev_ptr<float>(sv, sv->off_results + off);
    // the kernel reads the slot only through ring entries published after this
    std::memcpy(&sv->reqs[r->slot], &d, sizeof(d));
    sv->active.push_back(r);
  }
}

// Called with the lock held: count the ring entries the kernel has finished,
// hand finished requests to the copier, and move the frontier past them.
void harvest(bv_server* sv) {
  for (uint64_t e = sv->frontier; e < sv->published; ++e) {
    const uint32_t i = (uint32_t)(e & (kRingEntries - 1));
    if (sv->entry_done[i]) continue;
    if (__atomic_load_n(&sv->done[i], __ATOMIC_ACQUIRE) != (uint32_t)(e + 1)) continue;
    sv->entry_done[i] = 1;
    sv->outstanding -= 1;
    Request* r = sv->entry_req[i];
    r->done_tiles += 1;
    if (r->done_tiles == r->tiles) {
      sv->to_copy.push_back(r);
      sv->copy_cv.notify_one();
    }
  }
  while (sv->frontier < sv->published && sv->entry_done[sv->frontier & (kRingEntries - 1)]) sv->frontier += 1;
}

// Whether the running kernel can take r's tiles: r is of its variant and, for
// splats, r's scene's tree fits the kernel's LDS stack (a scene whose tree was
// built after the kernel started may be deeper).
bool servable(const bv_server* sv, const Request* r) {
  if (!sv->running || r->variant != sv->run_variant) return false;
  if (r->req.renderer != BV_RENDER_SPLAT) return true;
  return bvh::sphere_bvh_view(sv->scenes[r->req.scene]).stack_rows <= sv->run_rows;
}

// Called with the lock held: the variant's requests' next tiles into the ring,
// up to kQuantum from each request in turn, the requests with the fewest
// unpublished tiles first (shortest remaining work first: measured against
// arrival-order turns on the serving benchmark, 161 against 134 requests per
// second in a burst, the same p99), while fewer than tiles_ahead published
// tiles are unfinished (and the ring still holds the oldest unfinished one: a
// slow tile does not hold back the ones after it).
void publish(bv_server* sv) {
  std::vector<Request*> rs;
  for (Request* r : sv->active)
    if (servable(sv, r) && r->next_tile < r->tiles) rs.push_back(r);
  if (rs.empty()) return;
  // fewest unpublished tiles first (arrival order among equals)
  std::stable_sort(rs.begin(), rs.end(), [](const Request* a, const Request* b) {
    return a->tiles - a->next_tile < b->tiles - b->next_tile;
  });
  const uint64_t before = sv->published;
  size_t k = 0;
  size_t idle_turns = 0;
  while (sv->outstanding < sv->tiles_ahead && sv->published - sv->frontier < kRingEntries &&
         idle_turns < rs.size()) {
    Request* r = rs[k];
    k = (k + 1) % rs.size();
    if (r->next_tile >= r->tiles) {
      idle_turns += 1;
      continue;
    }
    idle_turns = 0;
    for (uint32_t q = 0; q < kQuantum && r->next_tile < r->tiles && sv->outstanding < sv->tiles_ahead &&
                         sv->published - sv->frontier < kRingEntries;
         ++q) {
      const uint32_t i = (uint32_t)(sv->published & (kRingEntries - 1));
      sv->ring[i] = ((uint32_t)r->slot << 20) | r->next_tile;
      sv->entry_req[i] = r;
      sv->entry_done[i] = 0;
      r->next_tile += 1;
      sv->published += 1;
      sv->outstanding += 1;
    }
  }
  if (sv->published != before) __atomic_store_n(&sv->ctl->published, sv->published, __ATOMIC_RELEASE);
}

// Called with the lock held: start the persistent kernel of `variant`.
void start_kernel(bv_server* sv, const Request* r) {
  const bv_scene* sc = sv->scenes[r->req.scene];
  const int mode = (int)r->req.mode;
  const bool splat = r->req.renderer == BV_RENDER_SPLAT;
  uint32_t rows = 2;
  int per_cu = 1;
  if (splat) {
    // the LDS stack covers the deepest tree among the server's scenes
    for (const bv_scene* s : sv->scenes)
      if (s->bounded && s->bvh_valid.load()) rows = std::max(rows, bvh::sphere_bvh_view(s).stack_rows);
    per_cu = bvh::serve_splat_blocks_per_cu(mode, sc->sh_degree, rows);
  } else {
    per_cu = bvh::serve_blocks_per_cu(mode, sc->sh_degree, sc->bounded != 0);
  }
  // one block slot per CU left free; block 0 mirrors the host's counters
  const uint32_t blocks = 1u + (uint32_t)sv->cu_count * (uint32_t)std::max(1, per_cu - 1);
  // enough queued tiles to refill every block between scheduler turns, few
  // enough that a new request's tiles start soon after it arrives
  sv->tiles_ahead = sv->cfg.tiles_ahead ? sv->cfg.tiles_ahead : blocks + blocks / 4;
  __atomic_store_n(&sv->ctl->stop, 0u, __ATOMIC_RELEASE);
  const bvk::ServeQueue q = queue_of(sv);
  hipLaunchKernelGGL(k_serve_reset, dim3(1), dim3(1), 0, sv->stream, q.claimed, q.published_dev, q.stop_dev,
                     sv->published);
  int rc = hipGetLastError() == hipSuccess ? BV_OK : BV_ERR_HIP;
  if (rc == BV_OK)
    rc = splat ? bvh::launch_serve_splat(mode, sc->sh_degree, q, blocks, rows, sv->stream)
               : bvh::launch_serve(mode, sc->sh_degree, sc->bounded != 0, q, blocks, sv->stream);
  if (rc != BV_OK || hipEventRecord(sv->kernel_done, sv->stream) != hipSuccess) {
    fail(sv, "serving kernel launch failed");
    return;
  }
  sv->running = true;
  sv->run_variant = r->variant;
  sv->run_rows = splat ? rows : 0u;
  sv->run_started = Clock::now();
}

// Called with the lock held (released while waiting): stop the running kernel
// once it has taken every published entry.
void stop_kernel(bv_server* sv, std::unique_lock<std::mutex>& lk) {
  __atomic_store_n(&sv->ctl->stop, 1u, __ATOMIC_RELEASE);
  lk.unlock();
  const hipError_t e = hipEventSynchronize(sv->kernel_done);
  lk.lock();
  sv->busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - sv->run_started).count();
  sv->running = false;
  sv->run_variant = -1;
  if (e != hipSuccess) fail(sv, std::string("serving kernel: ") + hipGetErrorString(e));
  harvest(sv);
}

void worker_loop(bv_server* sv) {
  (void)hipSetDevice(sv->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  sv->idle_since = Clock::now();
  while (true) {
    __atomic_store_n(&sv->ctl->heartbeat, ++sv->beat, __ATOMIC_RELEASE);
    admit(sv);
    harvest(sv);
    // the running variant: stop it when it has nothing left and another
    // variant waits, or after a spell without work
    if (sv->running) {
      bool work = false, other = false;
      for (const Request* r : sv->active) {
        if (r->done_tiles == r->tiles) continue;
        if (servable(sv, r))
          work = true;
        else
          other = true;
      }
      const hipError_t q = hipEventQuery(sv->kernel_done);
      if (q == hipSuccess) {
        fail(sv, "serving kernel left before it was stopped");
      } else if (q != hipErrorNotReady) {
        fail(sv, std::string("serving kernel: ") + hipGetErrorString(q));
      } else if (!work) {
        if (sv->idle) {
          const auto idle_us = std::chrono::duration_cast<std::chrono::microseconds>(Clock::now() - sv->idle_since);
          if (other || sv->stop || idle_us.count() > kIdleStopUs) stop_kernel(sv, lk);
        } else {
          sv->idle = true;
          sv->idle_since = Clock::now();
          if (other) stop_kernel(sv, lk);
        }
      } else {
        sv->idle = false;
      }
    }
    if (!sv->running && sv->error == BV_OK) {
      // the oldest admitted request with tiles left picks the variant
      const Request* first = nullptr;
      for (const Request* r : sv->active)
        if (r->done_tiles < r->tiles && (!first || r->ticket < first->ticket)) first = r;
      if (first) {
        start_kernel(sv, first);
        sv->idle = false;
      }
    }
    if (sv->running) publish(sv);
    sv->cv.notify_all();
    if (sv->error != BV_OK) {
      if (sv->running) stop_kernel(sv, lk);
      // fail every waiting request rather than hang clients
      for (auto& kv : sv->all) kv.second->complete = true;
      sv->cv.notify_all();
      break;
    }
    if (sv->stop && sv->pending.empty() && sv->active.empty() && !sv->running) break;
    if (sv->running && (sv->outstanding > 0 || !sv->idle)) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(kTurnUs));
      lk.lock();
    } else {
      sv->cv.wait_for(lk, std::chrono::microseconds(sv->running ? kIdleStopUs / 4 : 50000));
    }
  }
}

// Copies finished images from the result arena to the clients' buffers, then
// completes the requests.
void copier_loop(bv_server* sv) {
  std::unique_lock<std::mutex> lk(sv->mu);
  while (true) {
    sv->copy_cv.wait(lk, [&] { return !sv->to_copy.empty() || sv->copier_stop; });
    if (sv->to_copy.empty()) break;
    Request* r = sv->to_copy.front();
    sv->to_copy.pop_front();
    lk.unlock();
    std::memcpy(r->req.host_rgba, sv->results + r->off, (size_t)r->req.width * r->req.height * 16);
    lk.lock();
    sv->arena.release(r->off, r->bytes);
    sv->free_slots.push_back(r->slot);
    sv->active.erase(std::find(sv->active.begin(), sv->active.end(), r));
    r->complete = true;
    const auto now = Clock::now();
    r->latency_ms = std::chrono::duration<double, std::milli>(now - r->submitted).count();
    sv->latencies.push_back(r->latency_ms);
    sv->completed += 1;
    sv->rays_done += (uint64_t)r->req.width * r->req.height;
    sv->last_complete = now;
    sv->cv.notify_all();
  }
}

}  // namespace

extern "C" {

int bv_server_create(const bv_server_config* cfg, bv_server** out) {
  BV_REQUIRE(cfg && out, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(cfg->hbm_budget_bytes >= (1 << 20), BV_ERR_INVALID, "budget below 1 MiB");
  BV_REQUIRE(cfg->max_active >= 1 && cfg->max_active <= 4095, BV_ERR_INVALID, "max_active must be 1..4095");
  BV_REQUIRE(cfg->tiles_ahead <= kRingEntries, BV_ERR_INVALID, "tiles_ahead exceeds the ring");
  *out = nullptr;
  bv_server* sv = new bv_server();
  sv->cfg = *cfg;
  auto bail = [&](int code, const char* msg) {
    bvh::set_error(msg);
    if (sv->d_counters) (void)hipFree(sv->d_counters);
    if (sv->pinned) (void)hipHostFree(sv->pinned);
    if (sv->kernel_done) (void)hipEventDestroy(sv->kernel_done);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    delete sv;
    return code;
  };
  if (hipGetDevice(&sv->device) != hipSuccess) return bail(BV_ERR_HIP, "hipGetDevice failed");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, sv->device) != hipSuccess) return bail(BV_ERR_HIP, "device query failed");
  sv->cu_count = prop.multiProcessorCount;
  int wall_khz = 0;
  if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, sv->device) != hipSuccess || wall_khz <= 0)
    wall_khz = 100000;
  sv->dead_ticks = (uint64_t)(kHeartbeatTimeoutS * 1000.0 * wall_khz);
  if (hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&sv->kernel_done, hipEventDisableTiming) != hipSuccess)
    return bail(BV_ERR_HIP, "stream or event creation failed");
  // pinned host memory the kernel reads and writes: control block, ring,
  // completion words, request table, result arena
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  sv->off_ring = up(sizeof(bvk::ServeCtl));
  sv->off_done = up(sv->off_ring + sizeof(uint32_t) * kRingEntries);
  sv->off_reqs = up(sv->off_done + sizeof(uint32_t) * kRingEntries);
  sv->off_results = up(sv->off_reqs + sizeof(bvk::ServeReq) * cfg->max_active);
  const size_t pinned_bytes = sv->off_results + (size_t)cfg->hbm_budget_bytes;
  if (hipHostMalloc(&sv->pinned, pinned_bytes, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
    sv->pinned = nullptr;
    return bail(BV_ERR_OOM, "pinned result arena allocation failed");
  }
  void* dbase = nullptr;
  if (hipHostGetDevicePointer(&dbase, sv->pinned, 0) != hipSuccess)
    return bail(BV_ERR_HIP, "no device address for the pinned arena");
  sv->device_base = static_cast<char*>(dbase);
  char* base = static_cast<char*>(sv->pinned);
  std::memset(base, 0, sv->off_results);
  sv->ctl = reinterpret_cast<bvk::ServeCtl*>(base);
  sv->ring = reinterpret_cast<uint32_t*>(base + sv->off_ring);
  sv->done = reinterpret_cast<uint32_t*>(base + sv->off_done);
  sv->reqs = reinterpret_cast<bvk::ServeReq*>(base + sv->off_reqs);
  sv->results = base + sv->off_results;
  sv->arena.init(cfg->hbm_budget_bytes);
  if (hipMalloc(&sv->d_counters, 3 * sizeof(uint64_t)) != hipSuccess) {
    sv->d_counters = nullptr;
    return bail(BV_ERR_OOM, "device counter allocation failed");
  }
  sv->entry_req.assign(kRingEntries, nullptr);
  sv->entry_done.assign(kRingEntries, 1);
  for (int s = (int)cfg->max_active - 1; s >= 0; --s) sv->free_slots.push_back(s);
  sv->worker = std::thread(worker_loop, sv);
  sv->copier = std::thread(copier_loop, sv);
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
    if (!sc->bvh_valid.load()) {
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
  // the ticket is released: no thread refers to a completed request but
  // through `all` (its ring entries are done, its slot and arena range freed)
  sv->all.erase(it);
  delete r;
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
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->copier_stop = true;
  }
  sv->copy_cv.notify_all();
  if (sv->copier.joinable()) sv->copier.join();
  (void)hipStreamSynchronize(sv->stream);
  for (auto& kv : sv->all) delete kv.second;
  if (sv->d_counters) (void)hipFree(sv->d_counters);
  if (sv->pinned) (void)hipHostFree(sv->pinned);
  if (sv->kernel_done) (void)hipEventDestroy(sv->kernel_done);
  if (sv->stream) (void)hipStreamDestroy(sv->stream);
  delete sv;
  return BV_OK;
}

}  // extern "C"
