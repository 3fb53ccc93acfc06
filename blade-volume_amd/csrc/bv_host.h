// Host-side scene object and error plumbing shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "bv_internal.h"

struct bv_scene {
  int device = 0;
  int64_t n = 0;
  int64_t n_edges = 0;
  int32_t sh_degree = 0;
  int32_t K = 1;
  int32_t attr_dim = 4;
  int32_t bounded = 0;
  uint32_t max_degree = 0;
  int64_t device_bytes = 0;

  // reference layout (walked by the plain port)
  float4* pts = nullptr;          // [n] xyz + radius
  float* attrs = nullptr;         // [n][3K+1]
  uint32_t* offsets = nullptr;    // [n+1]
  uint32_t* neighbors = nullptr;  // [E]
  // walk layout
  bvk::CellRec* cells = nullptr;  // [n]
  bvk::EdgeRec* erec = nullptr;   // [E + 1] (the last record is zero padding)
  float* shd = nullptr;           // [n][K] float4 colour rows
  float4* material = nullptr;     // [n] albedo rgb + roughness (relighting), or null
  int32_t shd_stride = 4;
  // parameter masters
  float* xyz = nullptr;           // [n][3]
  float* density = nullptr;       // [n]
  float* sh_raw = nullptr;        // [n][3K]
  float* radius = nullptr;        // [n] (power cells only)
  // sphere BVH of the splat tracer (power cells), built on first use and
  // invalidated by every geometry rebuild
  void* bvh_nodes = nullptr;      // [max(n-1,1)] 64-byte inner nodes
  void* bvh_leaf = nullptr;       // [n] cell records in Morton order
  uint32_t* bvh_cell = nullptr;   // [n] site index per leaf
  // set (after bvh_depth and the arrays) once the tree is built: the serving
  // scheduler reads it on its own thread while a client's call may be building
  std::atomic<int> bvh_valid{0};
  uint32_t bvh_depth = 0;         // longest root-to-leaf path of the splat BVH (inner nodes)
  // Per-stream work buffers, so training calls on different streams never
  // share rows: the fused kernel's path cache (the forward's path per ray,
  // replayed by its adjoint) and the plain pipeline's rgba + dL/drgba.
  struct StreamBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<hipStream_t, StreamBuf> path_bufs;
  std::map<hipStream_t, StreamBuf> scratch_bufs;
  // the containing-cell query's 8-byte reduction slot, per stream, so queries
  // of one scene on different streams never share it
  std::map<hipStream_t, StreamBuf> argmin_bufs;
  // the persistent trace's lot counter, per stream
  std::map<hipStream_t, StreamBuf> counter_bufs;
  std::mutex path_mu;

  bvk::SceneView view() const {
    bvk::SceneView v;
    v.pts = pts;
    v.attrs = attrs;
    v.offsets = offsets;
    v.neighbors = neighbors;
    v.cells = cells;
    v.n = n;
    v.sh_degree = sh_degree;
    v.K = K;
    v.attr_dim = attr_dim;
    v.erec = erec;
    v.shd = shd;
    v.material = material;
    v.shd_stride = shd_stride;
    v.bounded = bounded;
    return v;
  }
};

// A CSR adjacency: on the device (bv_compute_cech) or on the host
// (bv_compute_delaunay, whose builder runs on the host).
struct bv_adjacency {
  int64_t n = 0;
  uint64_t n_edges = 0;
  uint32_t* offsets = nullptr;    // device [n + 1]
  uint32_t* neighbors = nullptr;  // device [n_edges]
  std::vector<uint32_t> host_offsets, host_neighbors;  // host copies when built there
};

namespace bvh {

void set_error(const std::string& msg);

// Returns from the calling C-ABI function with BV_ERR_HIP when a HIP call fails.
#define BV_HIP(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      bvh::set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));                   \
      return BV_ERR_HIP;                                                                          \
    }                                                                                             \
  } while (0)

#define BV_REQUIRE(cond, code, msg)   \
  do {                                \
    if (!(cond)) {                    \
      bvh::set_error(msg);            \
      return (code);                  \
    }                                 \
  } while (0)

// Rebuild the fast layout and the reference layout from the parameter masters.
int rebuild_layouts(bv_scene* s, hipStream_t stream);

// Fill a RaySource from a bv_rays batch.
bvk::RaySource ray_source(const bv_rays* rays);
// The same for a walk of scene s: per-ray start cells outside the scene make
// their rays empty (fetch_ray).
bvk::RaySource walk_rays(const bv_scene* s, const bv_rays* rays);
// Fill a RaySource for a camera image; tan_half is evaluated once on the host.
bvk::RaySource camera_source(const bv_camera* cam, uint32_t width, uint32_t height, uint32_t start_cell);
bvk::Settings settings_of(const bv_trace_settings* s);

// Cooperative walk launchers (bv_walk.hip).  lanes = lanes per ray (0 = choose).
int launch_trace_coop(bv_scene* s, const bvk::RaySource& rs, const bvk::Settings& st, const bvk::Outputs& out,
                      int lanes, hipStream_t stream);
int launch_record_coop(const bv_scene* s, const bvk::RaySource& rs, uint32_t L, float depth, float max_dt,
                       uint32_t* cells, uint32_t* next_cells, float* dts, float* mask, uint32_t* status, int lanes,
                       hipStream_t stream);
int launch_record_coop(const bv_scene* s, const bvk::ViewRays& rs, uint32_t L, float depth, float max_dt,
                       uint32_t* cells, uint32_t* next_cells, float* dts, float* mask, uint32_t* status, int lanes,
                       hipStream_t stream);
// PowerFoam recording with Jacobian streams over camera slices (bv_splat.hip);
// the sphere hierarchy must already exist (check_splat_scene).
int launch_splat_record_views(const bv_scene* s, const bvk::ViewRays& rs, uint32_t L, float depth, float max_dt,
                              const bv_splat_path_outputs& out, hipStream_t stream);
int check_splat_scene(bv_scene* s, hipStream_t stream);
// Multi-object scene walk (bv_walk.hip); objects share sh_degree and boundedness.
int launch_trace_objects(const bvk::ObjectSet& os, int32_t sh_degree, bool bounded, const float* origins,
                         const float* dirs, int64_t count, const bvk::Settings& st, float* rgba, hipStream_t stream);
struct GradPtrs {
  float* d_density;
  float* d_sh;
  float* d_xyz;
  // When set, the fused training replay adds each cell's density and position
  // gradient into one 16 B slot (x, y, z, density) here, added to d_density /
  // d_xyz after the launch: one memory atomic line per segment instead of two.
  float4* packed = nullptr;
};
int launch_adjoint_coop(bv_scene* s, const bvk::RaySource& rs, const bvk::Settings& st, const float* rgba_fwd,
                        const float* d_rgba, const GradPtrs& g, int lanes, hipStream_t stream);
int launch_train_l2_coop(bv_scene* s, const bvk::RaySource& rs, const bvk::Settings& st, const float* target,
                         float* rgba_out, float* loss, const GradPtrs& g, int lanes, hipStream_t stream,
                         uint32_t* path = nullptr, uint32_t path_len = 0);
// The fused training kernel over the library's fixed device workspace
// (kNoAllocScratchBytes per device, shared by every call on that device): no
// allocation, no packed gradient table.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
int launch_train_l2_noalloc(const bv_scene* s, const bvk::RaySource& rs, const bvk::Settings& st,
                            const float* target, float* rgba_out, float* loss, const GradPtrs& g,
                            hipStream_t stream);
int launch_relight_coop(const bv_scene* s, const bvk::RaySource& rs, const bvk::Settings& st,
                        const bv_lighting& lit, float* rgba_out, int lanes, hipStream_t stream);
// Serving: the persistent kernel of one walk variant (mode, SH degree,
// bounded) over the scheduler's ring, and the blocks of it one CU holds.
int launch_serve(int mode, int sh_degree, bool bounded, const bvk::ServeQueue& q, uint32_t blocks,
                 hipStream_t stream);
int serve_blocks_per_cu(int mode, int sh_degree, bool bounded);
// Linear BVH of n items (bv_bvh.hip): centres[i].xyz order the items, box_lo /
// box_hi [n] (original order) are their padded boxes; fills nodes
// [max(n-1,1)] and perm [n] (Morton position -> item).  Synchronous on stream.
// Fails (BV_ERR_UNSUPPORTED) if the tree is deeper than the traversal stack
// holds; *depth_out (may be null) receives the longest root-to-leaf path in
// inner nodes.
int build_lbvh(int64_t n, const float4* centres, const float4* box_lo, const float4* box_hi, bvk::BvhNode* nodes,
               uint32_t* perm, hipStream_t stream, uint32_t* depth_out = nullptr);
// Sphere BVH of a power-cell scene (bv_splat.hip); built once per geometry.
int ensure_sphere_bvh(bv_scene* s, hipStream_t stream);
bvk::BvhView sphere_bvh_view(const bv_scene* s);
// Serving: the persistent splat kernel (mode SH or RELIGHT) over the
// scheduler's ring; stack_rows covers the deepest BVH among the server's
// scenes.
int launch_serve_splat(int mode, int sh_degree, const bvk::ServeQueue& q, uint32_t blocks, uint32_t stack_rows,
                       hipStream_t stream);
int serve_splat_blocks_per_cu(int mode, int sh_degree, uint32_t stack_rows);
// A multiplier m with x -> x m mod n a bijection of [0, n) that scatters
// neighbours (the first prime from about 0.618 n on that does not divide n):
// kernels take their blocks of rays in that order, so rays that the caller
// grouped by camera (and that cross the same cells) do not run side by side
// and meet in their gradient atomics.
uint32_t scatter_mult(uint32_t n);
// The stream's buffer of at least `bytes` from `bufs` (grown after the stream's
// earlier work has finished with it).
int stream_buffer(bv_scene* s, std::map<hipStream_t, bv_scene::StreamBuf>& bufs, hipStream_t stream, size_t bytes,
                  void** out);
// Lanes per ray: BV_LANES_PER_RAY overrides the size-based choice.
int lanes_per_ray(int64_t rays);
// Steps per ray the fused training kernel's path cache holds (BV_PATH_CACHE_STEPS overrides).
uint32_t path_cache_steps();
int check_rays(const bv_scene* s, const bv_rays* rays);
// Delaunay edge graph of n FP32 sites (host arrays), rows sorted and self-free.
int delaunay_csr(const float* xyz, int64_t n, std::vector<uint32_t>* offsets, std::vector<uint32_t>* neighbors);

}  // namespace bvh
