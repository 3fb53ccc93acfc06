// Forward walk kernels (plain port and fast layout), camera rays and the path recorder.
//
// Walk semantics are trace_one_ray (blade-volume/src/trace.rs:1680-1820): the
// loop counts an iteration before testing `transmittance <= weight_threshold ||
// t0 > depth`, integrates the (support-clipped) cell interval when density >
// 1e-6, and stops when no exit face exists.  The plain kernel walks the
// reference's own device layout the way radfoam_trace.wgsl:104-219 does (one
// dependent gather of points[neighbour] per face); the fast kernel walks the
// per-edge layout (one contiguous 16-byte record per face, no dependent gather).
// Both evaluate each plane test with the same FP32 operations, so they visit the
// same cells as the oracle bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "bv_host.h"

namespace {

using namespace bvk;

// ---------------------------------------------------------------- plain port
// One thread per ray; every face gathers points[neighbour] (float4 xyz + radius)
// and recomputes the radical-plane shift, exactly the access pattern of
// radfoam_trace.wgsl:126-154.  Colour reads the [3K SH | density] attribute row.
template <int EVAL>
__global__ void __launch_bounds__(64) k_trace_plain(SceneView s, RaySource rs, Settings st, Outputs out) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= ray_slots(rs)) return;
  V3 ro, dir;
  uint32_t cur;
  int64_t oi;
  if (!fetch_ray(rs, slot, &ro, &dir, &cur, &oi)) return;
  WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
  if (!normalise(&dir)) {
    store_out(out, oi, a, 0u, cur, 0.0f);
    return;
  }
  float4 p = s.pts[cur];
  V3 cp = mk(p.x, p.y, p.z);
  float cr = p.w;
  float t0 = 0.0f;
  uint32_t steps = 0;
  const bool bounded = s.bounded != 0;
  while (steps < st.max_steps) {
    steps += 1;
    if (out.path != nullptr && steps - 1 < out.path_len) out.path[oi * (int64_t)out.path_len + steps - 1] = cur;
    if (a.trans <= st.weight_threshold || t0 > st.depth) break;
    const uint32_t b = s.offsets[cur], e = s.offsets[cur + 1];
    if (out.stats) a.bytes += step_bytes(e - b, bounded);
    float t1 = st.depth;
    uint32_t nb_best = BV_NO_CELL;
    const float ri_sq = cr * cr;
    for (uint32_t j = b; j < e; ++j) {
      const uint32_t nb = s.neighbors[j];
      const float4 q = s.pts[nb];
      const V3 off = sub(mk(q.x, q.y, q.z), cp);
      // radfoam_trace.wgsl:140-143 forms the radical-plane shift for every
      // face, Voronoi or not (radius 0 makes it exactly 0.5).
      float dsq = dot(off, off);
      dsq = dsq > 1e-20f ? dsq : 1e-20f;
      const float shift = 0.5f + 0.5f * (ri_sq - q.w * q.w) / dsq;
      plane_test(cp, off, shift, ro, dir, t0, nb, &t1, &nb_best);
    }
    float seg0, seg1;
    if (support_interval(bounded, ro, dir, cp, cr, t0, t1, &seg0, &seg1)) {
      const float* row = s.attrs + (int64_t)cur * s.attr_dim;
      const float sd = row[3 * s.K];
      if (sd > 1e-6f) {
        V3 col = mk(0.f, 0.f, 0.f);
        if (EVAL == BV_EVAL_SH) col = clamp0(eval_sh_dyn(s.sh_degree, row, dir));
        if (EVAL == BV_EVAL_SH && out.stats) a.bytes += 12u * (uint32_t)s.K;
        integrate<EVAL>(st, sd, seg0, seg1, cur, col, &a);
      }
    }
    if (nb_best == BV_NO_CELL) {
      t0 = t1;
      break;
    }
    t0 = t0 > t1 ? t0 : t1;
    cur = nb_best;
    p = s.pts[cur];
    cp = mk(p.x, p.y, p.z);
    cr = p.w;
  }
  store_out(out, oi, a, steps, cur, t0);
}

__global__ void k_camera_rays(RaySource rs, float* o, float* d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rs.width * rs.height) return;
  const uint32_t x = (uint32_t)(i % rs.width), y = (uint32_t)(i / rs.width);
  V3 ro, rd;
  camera_ray(rs, x, y, &ro, &rd);
  o[3 * i] = ro.x; o[3 * i + 1] = ro.y; o[3 * i + 2] = ro.z;
  d[3 * i] = rd.x; d[3 * i + 1] = rd.y; d[3 * i + 2] = rd.z;
}

int launch_trace(bv_scene* s, const RaySource& rs, const bv_trace_settings* settings, const bv_trace_outputs* o,
                 int kernel, hipStream_t stream) {
  BV_REQUIRE(settings != nullptr && o != nullptr, BV_ERR_INVALID, "null settings/outputs");
  BV_REQUIRE(settings->eval_mode <= BV_EVAL_SH, BV_ERR_INVALID, "bad eval_mode");
  BV_REQUIRE(kernel == BV_KERNEL_PLAIN || kernel == BV_KERNEL_FAST, BV_ERR_INVALID, "bad kernel id");
  const int64_t slots = ray_slots(rs);
  if (slots == 0) return BV_OK;
  const Settings st = bvh::settings_of(settings);
  Outputs out{o->rgba,       o->steps,      o->last_point, o->t_end, o->depth_mode,
              o->peak_weight, o->peak_point, o->stats,      o->path,  o->path ? o->path_len : 0u};
  const SceneView v = s->view();
  if (kernel == BV_KERNEL_PLAIN) {
    const int bs = 64;
    const dim3 grid((unsigned)((slots + bs - 1) / bs));
    switch (settings->eval_mode) {
      case BV_EVAL_OPACITY: hipLaunchKernelGGL((k_trace_plain<BV_EVAL_OPACITY>), grid, dim3(bs), 0, stream, v, rs, st, out); break;
      case BV_EVAL_CONSTANT: hipLaunchKernelGGL((k_trace_plain<BV_EVAL_CONSTANT>), grid, dim3(bs), 0, stream, v, rs, st, out); break;
      default: hipLaunchKernelGGL((k_trace_plain<BV_EVAL_SH>), grid, dim3(bs), 0, stream, v, rs, st, out); break;
    }
  } else {
    return bvh::launch_trace_coop(s, rs, st, out, 0, stream);
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int check_camera(const bv_camera* cam, uint32_t width, uint32_t height) {
  BV_REQUIRE(cam != nullptr, BV_ERR_INVALID, "camera is null");
  BV_REQUIRE(width > 0 && height > 0, BV_ERR_INVALID, "image size must be positive");
  BV_REQUIRE((uint64_t)width * height < (1ull << 40), BV_ERR_INVALID, "image too large");
  for (float f : cam->position) BV_REQUIRE(std::isfinite(f), BV_ERR_INVALID, "camera position must be finite");
  for (float f : cam->orientation) BV_REQUIRE(std::isfinite(f), BV_ERR_INVALID, "camera orientation must be finite");
  return BV_OK;
}

}  // namespace

extern "C" {

int bv_trace_rays(bv_scene* s, const bv_rays* rays, const bv_trace_settings* settings, const bv_trace_outputs* outputs,
                  int kernel, void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  return launch_trace(s, bvh::walk_rays(s, rays), settings, outputs, kernel, (hipStream_t)stream);
}

int bv_trace_camera(bv_scene* s, const bv_camera* cam, uint32_t width, uint32_t height, uint32_t start_cell,
                    const bv_trace_settings* settings, const bv_trace_outputs* outputs, int kernel, void* stream) {
  BV_REQUIRE(s != nullptr, BV_ERR_INVALID, "scene is null");
  int rc = check_camera(cam, width, height);
  if (rc != BV_OK) return rc;
  if (start_cell == BV_NO_CELL) {
    rc = bv_containing_cell(s, cam->position, &start_cell, stream);
    if (rc != BV_OK) return rc;
  }
  BV_REQUIRE((int64_t)start_cell < s->n, BV_ERR_INVALID, "start_cell out of bounds");
  return launch_trace(s, bvh::camera_source(cam, width, height, start_cell), settings, outputs, kernel,
                      (hipStream_t)stream);
}

int bv_trace_objects(const bv_scene_object* objects, uint32_t n_objects, const bv_rays* rays,
                     const bv_trace_settings* settings, float* out_rgba, void* stream) {
  BV_REQUIRE(rays != nullptr && settings != nullptr, BV_ERR_INVALID, "null rays/settings");
  BV_REQUIRE(n_objects <= BV_MAX_OBJECTS, BV_ERR_INVALID, "more than BV_MAX_OBJECTS objects");
  BV_REQUIRE(n_objects == 0 || objects != nullptr, BV_ERR_INVALID, "objects is null");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(settings->eval_mode <= BV_EVAL_SH, BV_ERR_INVALID, "bad eval_mode");
  BV_REQUIRE(settings->eval_mode != BV_EVAL_OPACITY, BV_ERR_UNSUPPORTED,
             "object scenes composite colour: eval_mode must be BV_EVAL_SH or BV_EVAL_CONSTANT");
  bvk::ObjectSet os{};
  os.n = n_objects;
  int32_t deg = 0;
  int32_t bounded = 0;
  for (uint32_t k = 0; k < n_objects; ++k) {
    const bv_scene_object& o = objects[k];
    BV_REQUIRE(o.scene != nullptr, BV_ERR_INVALID, "object scene is null");
    BV_REQUIRE((int64_t)o.start_cell < o.scene->n, BV_ERR_INVALID, "object start_cell out of bounds");
    bool finite = std::isfinite(o.radius) && o.radius >= 0.0f;
    for (int j = 0; j < 12; ++j) finite = finite && std::isfinite(o.world_to_object[j]);
    for (int j = 0; j < 3; ++j) finite = finite && std::isfinite(o.center[j]);
    BV_REQUIRE(finite, BV_ERR_INVALID, "object transform and bounds must be finite (radius >= 0)");
    const float* m = o.world_to_object;
    const double det = (double)m[0] * ((double)m[5] * m[10] - (double)m[6] * m[9]) -
                       (double)m[1] * ((double)m[4] * m[10] - (double)m[6] * m[8]) +
                       (double)m[2] * ((double)m[4] * m[9] - (double)m[5] * m[8]);
    BV_REQUIRE(std::fabs(det) > 1e-24, BV_ERR_INVALID, "object transform must be invertible");
    const bvk::SceneView v = o.scene->view();
    if (k == 0) {
      deg = v.sh_degree;
      bounded = v.bounded;
    }
    BV_REQUIRE(v.sh_degree == deg && v.bounded == bounded, BV_ERR_UNSUPPORTED,
               "objects of one call must share SH degree and RadFoam/PowerFoam kind");
    os.o[k].v = v;
    for (int j = 0; j < 12; ++j) os.o[k].m[j] = m[j];
    for (int j = 0; j < 3; ++j) os.o[k].center[j] = o.center[j];
    os.o[k].radius = o.radius;
    os.o[k].start = o.start_cell;
  }
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rays->origins != nullptr && rays->directions != nullptr && out_rgba != nullptr, BV_ERR_INVALID,
             "null ray or output buffer");
  return bvh::launch_trace_objects(os, deg, bounded != 0, rays->origins, rays->directions, rays->count,
                                   bvh::settings_of(settings), out_rgba, (hipStream_t)stream);
}

int bv_camera_rays(const bv_camera* cam, uint32_t width, uint32_t height, float* origins, float* directions,
                   void* stream) {
  int rc = check_camera(cam, width, height);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(origins && directions, BV_ERR_INVALID, "null output");
  const bvk::RaySource rs = bvh::camera_source(cam, width, height, 0);
  const int64_t n = (int64_t)width * height;
  hipLaunchKernelGGL(k_camera_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, rs, origins,
                     directions);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

static int check_lighting(const bv_scene* s, const bv_lighting* lit) {
  BV_REQUIRE(lit != nullptr, BV_ERR_INVALID, "lighting is null");
  BV_REQUIRE(lit->count >= 1 && lit->count <= BV_MAX_LIGHTS, BV_ERR_INVALID, "lighting needs 1..8 lights");
  BV_REQUIRE(s->material != nullptr, BV_ERR_INVALID, "relighting needs a scene created with materials");
  return BV_OK;
}

int bv_relight_camera(bv_scene* s, const bv_camera* cam, uint32_t width, uint32_t height, uint32_t start_cell,
                      const bv_trace_settings* settings, const bv_lighting* lighting, float* rgba_out,
                      void* stream) {
  BV_REQUIRE(s != nullptr && settings != nullptr && rgba_out != nullptr, BV_ERR_INVALID, "null argument");
  int rc = check_camera(cam, width, height);
  if (rc != BV_OK) return rc;
  rc = check_lighting(s, lighting);
  if (rc != BV_OK) return rc;
  if (start_cell == BV_NO_CELL) {
    rc = bv_containing_cell(s, cam->position, &start_cell, stream);
    if (rc != BV_OK) return rc;
  }
  BV_REQUIRE((int64_t)start_cell < s->n, BV_ERR_INVALID, "start_cell out of bounds");
  return bvh::launch_relight_coop(s, bvh::camera_source(cam, width, height, start_cell), bvh::settings_of(settings),
                                  *lighting, rgba_out, 0, (hipStream_t)stream);
}

int bv_relight_rays(bv_scene* s, const bv_rays* rays, const bv_trace_settings* settings, const bv_lighting* lighting,
                    float* rgba_out, void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(settings != nullptr, BV_ERR_INVALID, "null settings");
  rc = check_lighting(s, lighting);
  if (rc != BV_OK) return rc;
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rgba_out != nullptr, BV_ERR_INVALID, "null output");
  return bvh::launch_relight_coop(s, bvh::walk_rays(s, rays), bvh::settings_of(settings), *lighting, rgba_out, 0,
                                  (hipStream_t)stream);
}

int bv_record_paths(bv_scene* s, const bv_rays* rays, const bv_record_settings* settings, uint32_t* cells,
                    uint32_t* next_cells, float* dts, float* mask, uint32_t* status, void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(settings != nullptr, BV_ERR_INVALID, "null settings");
  BV_REQUIRE(settings->max_steps > 0, BV_ERR_INVALID, "max_steps must be positive");
  BV_REQUIRE(!s->bounded, BV_ERR_UNSUPPORTED,
             "power-cell path recording (support-clipped intervals with Jacobians) is not in this build");
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(cells && next_cells && dts && status, BV_ERR_INVALID, "null output");
  const bvk::RaySource rs = bvh::walk_rays(s, rays);
  return bvh::launch_record_coop(s, rs, settings->max_steps, settings->depth, settings->max_path_dt, cells,
                                 next_cells, dts, mask, status, 0, (hipStream_t)stream);
}

int bv_record_paths_batch(bv_scene* s, const bv_record_view* views, uint32_t n_views, const uint32_t* pixel_indices,
                          uint32_t capacity, uint32_t max_steps, const bv_record_batch_outputs* out, void* stream) {
  BV_REQUIRE(s != nullptr && out != nullptr, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(n_views == 0 || views != nullptr, BV_ERR_INVALID, "null views");
  BV_REQUIRE(max_steps > 0, BV_ERR_INVALID, "max_steps must be positive");
  // dispatch_batch (path_record.rs:529-538) and prepare_dispatch (:330-337)
  int64_t end = 0, rows = 0;
  for (uint32_t i = 0; i < n_views; ++i) {
    const bv_record_view& v = views[i];
    BV_REQUIRE(v.image_width > 0 && v.image_height > 0, BV_ERR_INVALID, "record view has an empty image");
    BV_REQUIRE(v.num_pixels <= capacity && v.pixel_offset <= capacity - v.num_pixels, BV_ERR_INVALID,
               "path dispatch exceeds pixel buffer capacity");
    BV_REQUIRE((int64_t)v.pixel_offset >= end, BV_ERR_INVALID,
               "batched path dispatch ranges must be ordered and disjoint");
    BV_REQUIRE(s->bounded || v.num_pixels == 0 || (int64_t)v.start_point < s->n, BV_ERR_INVALID,
               "start_point out of bounds");
    end = (int64_t)v.pixel_offset + v.num_pixels;
    rows += v.num_pixels;
  }
  if (rows == 0) return BV_OK;
  const bv_splat_path_outputs& r = out->rows;
  BV_REQUIRE(pixel_indices != nullptr, BV_ERR_INVALID, "null pixel_indices");
  BV_REQUIRE(r.cells && r.next_cells && r.dts && r.status, BV_ERR_INVALID, "null record buffers");
  hipStream_t st = (hipStream_t)stream;
  if (s->bounded) {
    BV_REQUIRE(r.previous_cells && r.dt_grad_previous && r.dt_grad_current && r.dt_grad_next &&
                   r.dt_reference_tangent,
               BV_ERR_INVALID, "power-cell recording needs previous_cells and the Jacobian streams");
    int rc = bvh::check_splat_scene(s, st);
    if (rc != BV_OK) return rc;
  }
  bvk::ViewRays vr{};
  vr.pixel_indices = pixel_indices;
  vr.ray_origins = out->ray_origins;
  vr.ray_dirs = out->ray_dirs;
  vr.view_index = out->view_index;
  float depth = 0.0f, max_dt = 0.0f;
  auto flush = [&]() -> int {
    if (vr.n == 0) return BV_OK;
    const int rc = s->bounded ? bvh::launch_splat_record_views(s, vr, max_steps, depth, max_dt, r, st)
                              : bvh::launch_record_coop(s, vr, max_steps, depth, max_dt, r.cells, r.next_cells,
                                                        r.dts, r.mask, r.status, 0, st);
    vr.n = 0;
    vr.count = 0;
    return rc;
  };
  for (uint32_t i = 0; i < n_views; ++i) {
    const bv_record_view& v = views[i];
    if (v.num_pixels == 0) continue;
    // a launch holds slices of one depth and dt cap (compared bitwise)
    const bool same = vr.n > 0 && std::memcmp(&depth, &v.depth, 4) == 0 && std::memcmp(&max_dt, &v.max_path_dt, 4) == 0;
    if (!same || vr.n == (uint32_t)bvk::kViewsPerLaunch) {
      const int rc = flush();
      if (rc != BV_OK) return rc;
      depth = v.depth;
      max_dt = v.max_path_dt;
    }
    bvk::ViewSlice& d = vr.v[vr.n];
    for (int k = 0; k < 3; ++k) d.pos[k] = v.camera.position[k];
    for (int k = 0; k < 4; ++k) d.quat[k] = v.camera.orientation[k];
    for (int k = 0; k < 2; ++k) {
      d.tan_half[k] = (float)std::tan((double)(0.5f * v.camera.fov[k]));
      d.principal[k] = v.camera.principal[k];
    }
    d.width = v.image_width;
    d.height = v.image_height;
    d.start = v.start_point;
    d.pixel_offset = v.pixel_offset;
    d.image_pixel_offset = v.image_pixel_offset;
    d.view = i;
    d.slot_begin = vr.count;
    vr.count += v.num_pixels;
    vr.n += 1;
  }
  return flush();
}

}  // extern "C"
