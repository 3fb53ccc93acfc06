// Scene upload, derived device layouts and the containing-cell query.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bv_host.h"

namespace bvh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

bvk::RaySource ray_source(const bv_rays* rays) {
  bvk::RaySource rs{};
  rs.origins = rays->origins;
  rs.dirs = rays->directions;
  rs.start_cells = rays->start_cells;
  rs.start_cell = rays->start_cell;
  rs.count = rays->count;
  return rs;
}

bvk::RaySource walk_rays(const bv_scene* s, const bv_rays* rays) {
  bvk::RaySource rs = ray_source(rays);
  rs.n_cells = (uint32_t)s->n;
  return rs;
}

bvk::RaySource camera_source(const bv_camera* cam, uint32_t width, uint32_t height, uint32_t start_cell) {
  bvk::RaySource rs{};
  rs.origins = nullptr;
  rs.dirs = nullptr;
  rs.start_cells = nullptr;
  rs.start_cell = start_cell;
  rs.count = (int64_t)width * height;
  for (int k = 0; k < 3; ++k) rs.cam_pos[k] = cam->position[k];
  for (int k = 0; k < 4; ++k) rs.quat[k] = cam->orientation[k];
  // tan of the FP32 half angle, evaluated in double and rounded once, so every
  // consumer (device kernels, host tools, the tests) derives identical rays.
  rs.tan_half[0] = (float)std::tan((double)(0.5f * cam->fov[0]));
  rs.tan_half[1] = (float)std::tan((double)(0.5f * cam->fov[1]));
  rs.principal[0] = cam->principal[0];
  rs.principal[1] = cam->principal[1];
  rs.width = width;
  rs.height = height;
  return rs;
}

bvk::Settings settings_of(const bv_trace_settings* s) {
  bvk::Settings o{};
  o.weight_threshold = s->weight_threshold;
  o.max_steps = s->max_steps;
  o.depth = s->depth;
  o.eval_mode = s->eval_mode;
  for (int k = 0; k < 3; ++k) o.const_rgb[k] = s->const_rgb[k];
  return o;
}

int check_rays(const bv_scene* s, const bv_rays* rays) {
  BV_REQUIRE(s != nullptr, BV_ERR_INVALID, "scene is null");
  BV_REQUIRE(rays != nullptr, BV_ERR_INVALID, "rays is null");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(rays->count <= (int64_t)UINT32_MAX, BV_ERR_INVALID, "more than 2^32 - 1 rays in one call");
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rays->origins != nullptr && rays->directions != nullptr, BV_ERR_INVALID,
             "ray origins/directions are null");
  BV_REQUIRE(rays->start_cells != nullptr || (int64_t)rays->start_cell < s->n, BV_ERR_INVALID,
             "start_cell out of bounds");
  return BV_OK;
}

}  // namespace bvh

namespace {

using bvk::CellRec;

// One thread per site: the reference point / attribute rows and the walk's
// cell record.
__global__ void k_build_sites(int64_t n, int K, int attr_dim, const float* __restrict__ xyz,
                              const float* __restrict__ density, const float* __restrict__ sh_raw,
                              const float* __restrict__ radius, const uint32_t* __restrict__ offsets, float4* pts,
                              float* attrs, CellRec* cells) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const float r = radius ? radius[i] : 0.0f;
  const float dens = density[i];
  pts[i] = make_float4(x, y, z, r);
  float* arow = attrs + i * attr_dim;
  const float* src = sh_raw + i * 3 * K;
  for (int k = 0; k < 3 * K; ++k) arow[k] = src[k];
  arow[3 * K] = dens;
  CellRec c;
  c.x = x; c.y = y; c.z = z; c.density = dens;
  c.begin = offsets[i];
  c.end = offsets[i + 1];
  c.radius = r;
  c.pad = 0u;
  cells[i] = c;
}

// One thread per site: the cooperative walk's edge records of its row and its
// colour row.
__global__ void k_build_erec(int64_t n, int K, int shd_stride, const uint32_t* __restrict__ offsets,
                             const uint32_t* __restrict__ nbrs, const float* __restrict__ xyz,
                             const float* __restrict__ radius, const float* __restrict__ density,
                             const float* __restrict__ sh_raw, bvk::EdgeRec* erec, float* shd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = offsets[i], e = offsets[i + 1];
  const float xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
  const float ri = radius ? radius[i] : 0.0f;
  const float ri_sq = ri * ri;
  for (uint32_t k = b; k < e; ++k) {
    const uint32_t j = nbrs[k];
    bvk::EdgeRec r;
    r.x = xyz[3 * (int64_t)j];
    r.y = xyz[3 * (int64_t)j + 1];
    r.z = xyz[3 * (int64_t)j + 2];
    r.nb = j;
    r.begin = offsets[j];
    r.end = offsets[j + 1];
    r.radius = radius ? radius[j] : 0.0f;
    if (radius) {
      const float ox = r.x - xi, oy = r.y - yi, oz = r.z - zi;
      float dsq = (ox * ox) + (oy * oy) + (oz * oz);
      dsq = dsq > 1e-20f ? dsq : 1e-20f;
      r.shift = 0.5f + 0.5f * (ri_sq - r.radius * r.radius) / dsq;
    } else {
      r.shift = 0.5f;
    }
    erec[k] = r;
  }
  // Component-major colour row: float4 (r, g, b, .) per SH component, the
  // density in component 0's fourth slot and zeros in the others.
  float4* row = reinterpret_cast<float4*>(shd + i * shd_stride);
  const float* src = sh_raw + i * 3 * K;
  for (int k = 0; k < K; ++k) row[k] = make_float4(src[3 * k], src[3 * k + 1], src[3 * k + 2], k == 0 ? density[i] : 0.0f);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// Exact (power-)nearest site: argmin over (ordered distance bits, index), so
// the lowest index wins ties as in lib.rs:534-537.
__global__ void k_argmin(int64_t n, const float4* __restrict__ pts, float qx, float qy, float qz,
                         unsigned long long* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long key = ~0ull;
  if (i < n) {
    const float4 p = pts[i];
    const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
    const float d = ((dx * dx) + (dy * dy)) + (dz * dz) - p.w * p.w;
    if (!isnan(d)) {
      uint32_t u = __float_as_uint(d);
      u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
      key = ((unsigned long long)u << 32) | (unsigned long long)(uint32_t)i;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(key, off, 64);
    key = o < key ? o : key;
  }
  if ((threadIdx.x & 63) == 0 && key != ~0ull) atomicMin(out, key);
}

int upload(void** dst, const void* src, size_t bytes, int64_t* total) {
  if (bytes == 0) {
    *dst = nullptr;
    return BV_OK;
  }
  hipError_t e = hipMalloc(dst, bytes);
  if (e != hipSuccess) {
    bvh::set_error(std::string("hipMalloc(") + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    return BV_ERR_OOM;
  }
  *total += (int64_t)bytes;
  if (src) {
    e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      bvh::set_error(std::string("hipMemcpy failed: ") + hipGetErrorString(e));
      return BV_ERR_HIP;
    }
  }
  return BV_OK;
}

void free_scene(bv_scene* s) {
  void* ptrs[] = {s->pts, s->attrs, s->offsets, s->neighbors, s->cells,
                  s->erec, s->shd, s->xyz, s->density, s->sh_raw, s->radius,
                  s->material, s->bvh_nodes, s->bvh_leaf, s->bvh_cell};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto* m : {&s->path_bufs, &s->scratch_bufs, &s->argmin_bufs, &s->counter_bufs})
    for (auto& kv : *m)
      if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  delete s;
}

}  // namespace

namespace bvh {

uint32_t scatter_mult(uint32_t n) {
  if (n <= 2) return 1;
  for (uint64_t m = (uint64_t)((double)n * 0.6180339887) | 1u;; m += 2) {
    bool prime = m > 2;
    for (uint64_t q = 3; q * q <= m && prime; q += 2) prime = m % q != 0;
    if (prime && n % m != 0) return (uint32_t)m;
  }
}

int stream_buffer(bv_scene* s, std::map<hipStream_t, bv_scene::StreamBuf>& bufs, hipStream_t stream, size_t bytes,
                  void** out) {
  std::lock_guard<std::mutex> lock(s->path_mu);
  bv_scene::StreamBuf& b = bufs[stream];
  if (b.bytes < bytes) {
    if (b.ptr) {
      BV_HIP(hipStreamSynchronize(stream));  // earlier launches on this stream may still use it
      BV_HIP(hipFree(b.ptr));
    }
    b.ptr = nullptr;
    b.bytes = 0;
    if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
      bvh::set_error("hipMalloc of a per-stream work buffer failed");
      return BV_ERR_OOM;
    }
    b.bytes = bytes;
  }
  *out = b.ptr;
  return BV_OK;
}

int rebuild_layouts(bv_scene* s, hipStream_t stream) {
  s->bvh_valid.store(0);
  if (s->n == 0) return BV_OK;
  const int bs = 256;
  const unsigned grid = (unsigned)((s->n + bs - 1) / bs);
  hipLaunchKernelGGL(k_build_sites, dim3(grid), dim3(bs), 0, stream, s->n, s->K, s->attr_dim, s->xyz, s->density,
                     s->sh_raw, s->radius, s->offsets, s->pts, s->attrs, s->cells);
  BV_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_build_erec, dim3(grid), dim3(bs), 0, stream, s->n, s->K, s->shd_stride, s->offsets,
                     s->neighbors, s->xyz, s->radius, s->density, s->sh_raw, s->erec, s->shd);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

}  // namespace bvh

extern "C" {

const char* bv_last_error(void) { return bvh::g_last_error.c_str(); }

int bv_version(void) { return 1; }

int bv_device_info(int* device_count, int* cu_count) {
  int n = 0;
  BV_HIP(hipGetDeviceCount(&n));
  if (device_count) *device_count = n;
  if (cu_count) {
    *cu_count = 0;
    if (n > 0) {
      hipDeviceProp_t p;
      BV_HIP(hipGetDeviceProperties(&p, 0));
      *cu_count = p.multiProcessorCount;
    }
  }
  return BV_OK;
}

int bv_scene_create(const bv_scene_desc* d, bv_scene** out) {
  BV_REQUIRE(d != nullptr && out != nullptr, BV_ERR_INVALID, "null argument");
  *out = nullptr;
  BV_REQUIRE(d->flags == 0, BV_ERR_UNSUPPORTED,
             "oriented / surface-detail / spherical-Voronoi / surface-colour models are not supported");
  BV_REQUIRE(d->n_sites > 0, BV_ERR_INVALID, "model has no points");
  BV_REQUIRE(d->n_sites < (int64_t)0xFFFFFFFF, BV_ERR_INVALID, "too many sites for u32 cell ids");
  BV_REQUIRE(d->sh_degree >= 0 && d->sh_degree <= 3, BV_ERR_INVALID, "sh_degree must be in 0..=3");
  BV_REQUIRE(d->xyz && d->density && d->sh && d->offsets, BV_ERR_INVALID, "missing model array");
  const int64_t n = d->n_sites;
  // CSR validation (adjacency.rs:721-797): monotone offsets, in-range, self-free neighbours.
  BV_REQUIRE(d->offsets[0] == 0, BV_ERR_INVALID, "offsets[0] must be 0");
  uint32_t max_deg = 0;
  for (int64_t i = 0; i < n; ++i) {
    BV_REQUIRE(d->offsets[i + 1] >= d->offsets[i], BV_ERR_INVALID, "offsets must be non-decreasing");
    const uint32_t deg = d->offsets[i + 1] - d->offsets[i];
    max_deg = deg > max_deg ? deg : max_deg;
  }
  const int64_t E = d->offsets[n];
  BV_REQUIRE(E == 0 || d->neighbors != nullptr, BV_ERR_INVALID, "missing neighbours");
  for (int64_t i = 0; i < n; ++i)
    for (uint32_t k = d->offsets[i]; k < d->offsets[i + 1]; ++k) {
      BV_REQUIRE(d->neighbors[k] < (uint32_t)n, BV_ERR_INVALID, "neighbour index out of bounds");
      BV_REQUIRE(d->neighbors[k] != (uint32_t)i, BV_ERR_INVALID, "self neighbour in adjacency");
    }
  for (int64_t i = 0; i < 3 * n; ++i)
    BV_REQUIRE(std::isfinite(d->xyz[i]), BV_ERR_INVALID, "non-finite site position");
  if (d->material)
    for (int64_t i = 0; i < 4 * n; ++i) BV_REQUIRE(std::isfinite(d->material[i]), BV_ERR_INVALID, "bad material");
  for (int64_t i = 0; i < n; ++i) {
    BV_REQUIRE(std::isfinite(d->density[i]), BV_ERR_INVALID, "non-finite density");
    if (d->radius) BV_REQUIRE(std::isfinite(d->radius[i]) && d->radius[i] >= 0.0f, BV_ERR_INVALID, "bad radius");
  }

  int device = 0;
  BV_HIP(hipGetDevice(&device));
  bv_scene* s = new bv_scene();
  s->device = device;
  s->n = n;
  s->n_edges = E;
  s->sh_degree = d->sh_degree;
  s->K = (d->sh_degree + 1) * (d->sh_degree + 1);
  s->attr_dim = 3 * s->K + 1;
  s->shd_stride = 4 * s->K;
  s->bounded = d->radius != nullptr;
  s->max_degree = max_deg;
  int rc = BV_OK;
  int64_t& tot = s->device_bytes;
#define UP(field, src, bytes)                                           \
  if (rc == BV_OK) rc = upload((void**)&s->field, src, (size_t)(bytes), &tot)
  UP(xyz, d->xyz, 12 * n);
  UP(density, d->density, 4 * n);
  UP(sh_raw, d->sh, (size_t)12 * s->K * n);
  if (s->bounded) UP(radius, d->radius, 4 * n);
  UP(offsets, d->offsets, 4 * (n + 1));
  UP(neighbors, d->neighbors, 4 * E);
  UP(pts, nullptr, 16 * n);
  UP(attrs, nullptr, (size_t)4 * s->attr_dim * n);
  UP(cells, nullptr, sizeof(CellRec) * n);
  UP(erec, nullptr, sizeof(bvk::EdgeRec) * (E + 1));  // + a zero padding record (index E)
  UP(shd, nullptr, (size_t)4 * s->shd_stride * n);
  if (d->material) UP(material, d->material, 16 * n);
#undef UP
  if (rc == BV_OK && hipMemset(s->erec + E, 0, sizeof(bvk::EdgeRec)) != hipSuccess) {
    bvh::set_error("clearing the padding edge record failed");
    rc = BV_ERR_HIP;
  }
  if (rc == BV_OK) rc = bvh::rebuild_layouts(s, nullptr);
  if (rc == BV_OK) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      bvh::set_error(std::string("scene build failed: ") + hipGetErrorString(e));
      rc = BV_ERR_HIP;
    }
  }
  if (rc != BV_OK) {
    free_scene(s);
    return rc;
  }
  *out = s;
  return BV_OK;
}

int bv_scene_destroy(bv_scene* s) {
  if (!s) return BV_OK;
  // the scene's memory is released whatever the device reports
  const hipError_t e = hipDeviceSynchronize();
  free_scene(s);
  BV_HIP(e);
  return BV_OK;
}

int bv_scene_get_info(const bv_scene* s, bv_scene_info* info) {
  BV_REQUIRE(s && info, BV_ERR_INVALID, "null argument");
  info->n_sites = s->n;
  info->n_edges = s->n_edges;
  info->sh_degree = s->sh_degree;
  info->bounded = s->bounded;
  info->max_degree = s->max_degree;
  info->pad = 0;
  info->device_bytes = s->device_bytes;
  return BV_OK;
}

int bv_scene_update(bv_scene* s, const float* xyz, const float* density, const float* sh, void* stream) {
  BV_REQUIRE(s != nullptr, BV_ERR_INVALID, "scene is null");
  hipStream_t st = (hipStream_t)stream;
  if (xyz) BV_HIP(hipMemcpyAsync(s->xyz, xyz, 12 * (size_t)s->n, hipMemcpyDeviceToDevice, st));
  if (density) BV_HIP(hipMemcpyAsync(s->density, density, 4 * (size_t)s->n, hipMemcpyDeviceToDevice, st));
  if (sh) BV_HIP(hipMemcpyAsync(s->sh_raw, sh, 12 * (size_t)s->K * s->n, hipMemcpyDeviceToDevice, st));
  return bvh::rebuild_layouts(s, st);
}

int bv_containing_cell(bv_scene* s, const float position[3], uint32_t* out_cell, void* stream) {
  BV_REQUIRE(s && position && out_cell, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(std::isfinite(position[0]) && std::isfinite(position[1]) && std::isfinite(position[2]),
             BV_ERR_INVALID, "query position must be finite");
  hipStream_t st = (hipStream_t)stream;
  void* slot_buf = nullptr;
  const int rc = bvh::stream_buffer(s, s->argmin_bufs, st, 8, &slot_buf);
  if (rc != BV_OK) return rc;
  unsigned long long* slot = static_cast<unsigned long long*>(slot_buf);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, slot, ~0ull);
  BV_HIP(hipGetLastError());
  const int bs = 256;
  hipLaunchKernelGGL(k_argmin, dim3((unsigned)((s->n + bs - 1) / bs)), dim3(bs), 0, st, s->n, s->pts, position[0],
                     position[1], position[2], slot);
  BV_HIP(hipGetLastError());
  unsigned long long key = 0;
  BV_HIP(hipMemcpyAsync(&key, slot, 8, hipMemcpyDeviceToHost, st));
  BV_HIP(hipStreamSynchronize(st));
  BV_REQUIRE(key != ~0ull, BV_ERR_INVALID, "no finite site distance");
  *out_cell = (uint32_t)(key & 0xFFFFFFFFull);
  return BV_OK;
}

}  // extern "C"
