// Device-side layouts and the arithmetic of one walk step, shared by the trace,
// record and adjoint kernels.  Compiled with -ffp-contract=off so every product
// and sum below rounds exactly as the reference's scalar FP32 code does
// (trace.rs:170-205: no fused multiply-add, glam dot = (x*x')+(y*y')+(z*z')).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bv.h"

namespace bvk {

// Per-site record of the fast layout: position, density and the CSR row bounds
// in one 32-byte, 32-byte-aligned block (a single cache line per visited cell).
struct __align__(32) CellRec {
  float x, y, z, density;
  uint32_t begin, end;
  float radius;
  uint32_t pad;
};

// Per directed edge i -> j of the cooperative walk, in CSR order: everything
// the walk needs about neighbour j, so a step is one contiguous read of the
// current cell's row.  Position, CSR row bounds, radius, and the radical-plane
// shift 0.5 + 0.5 (r_i^2 - r_j^2) / max(|p_j - p_i|^2, 1e-20) (0.5 for Voronoi
// cells).  p_j is the site's own FP32 position, so the walk forms p_j - p_i at
// run time exactly as the reference does.
struct __align__(32) EdgeRec {
  float x, y, z;
  uint32_t nb;
  uint32_t begin, end;
  float radius, shift;
};

// Everything a walk kernel reads.  `pts` / `attrs` / `offsets` / `neighbors`
// are the reference's own device layout (gpu/radfoam.rs:336-413: points with
// radius in .w, attribute rows [3K SH | density]); the plain port walks them.
// `cells` / `erec` / `shd` are the cooperative walk's layout: one 32-byte cell
// record per site, one 32-byte edge record per directed edge in CSR order and
// the colour rows as float4 per SH component.
struct SceneView {
  // reference layout
  const float4* __restrict__ pts;
  const float* __restrict__ attrs;
  const uint32_t* __restrict__ offsets;
  const uint32_t* __restrict__ neighbors;
  // fast layout
  const CellRec* __restrict__ cells;
  const EdgeRec* __restrict__ erec;
  const float* __restrict__ shd;  // [n][K] float4 (r, g, b, density in component 0 else 0)
  const float4* __restrict__ material;  // [n] albedo rgb + roughness, or null
  int64_t n;
  int32_t sh_degree;
  int32_t K;          // SH components
  int32_t attr_dim;   // 3K + 1
  int32_t shd_stride; // floats per colour row (4K)
  int32_t bounded;
};

struct RaySource {
  const float* __restrict__ origins;
  const float* __restrict__ dirs;
  const uint32_t* __restrict__ start_cells;
  uint32_t start_cell;
  // Cell count of the walked scene when per-ray start cells come from the
  // caller (0: no per-ray check, the start cell was validated on the host).
  uint32_t n_cells;
  int64_t count;
  // camera mode (origins == nullptr)
  float cam_pos[3];
  float quat[4];
  float tan_half[2];
  float principal[2];
  uint32_t width, height;
};

struct Settings {
  float weight_threshold;
  uint32_t max_steps;
  float depth;
  uint32_t eval_mode;
  float const_rgb[3];
};

// One object of a multi-object scene (bv_scene_object in device form).
struct ObjectView {
  SceneView v;
  float m[12];  // world -> object, row-major 3x4
  float center[3];
  float radius;
  uint32_t start;
};
struct ObjectSet {
  ObjectView o[BV_MAX_OBJECTS];
  uint32_t n;
};

struct Outputs {
  float* rgba;
  uint32_t* steps;
  uint32_t* last_point;
  float* t_end;
  float* depth_mode;
  float* peak_weight;
  uint32_t* peak_point;
  uint32_t* stats;
  uint32_t* path;     // [count][path_len] current cell of each walk iteration, or null
  uint32_t path_len;
};

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 scl(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// SH constants of trace.rs:242-261 (f32 literals).
__constant__ static const float kShC[16] = {
    0.2820948f,  -0.48860252f, 0.48860252f, -0.48860252f, 1.0925485f, -1.0925485f,
    0.31539157f, -1.0925485f,  0.54627424f, -0.5900436f,  2.8906114f, -0.4570458f,
    0.37317634f, -0.4570458f,  1.4453057f,  -0.5900436f};

// eval_rgb_sh_unclamped (trace.rs:264-392) on a coefficient row `c` with the
// same term order: each term is (C_k * coef) * factor, then 0.5 + colour.
template <int DEG>
__device__ __forceinline__ V3 eval_sh_row(const float* __restrict__ c, V3 d) {
  float r = 0.f, g = 0.f, b = 0.f;
  r += kShC[0] * c[0];
  g += kShC[0] * c[1];
  b += kShC[0] * c[2];
  if (DEG >= 1) {
    const float f[3] = {d.y, d.z, d.x};
#pragma unroll
    for (int k = 1; k <= 3; ++k) {
      r += kShC[k] * c[3 * k] * f[k - 1];
      g += kShC[k] * c[3 * k + 1] * f[k - 1];
      b += kShC[k] * c[3 * k + 2] * f[k - 1];
    }
  }
  if (DEG >= 2) {
    const float x = d.x, y = d.y, z = d.z;
    const float xx = x * x, yy = y * y, zz = z * z;
    r += kShC[4] * c[12] * x * y; g += kShC[4] * c[13] * x * y; b += kShC[4] * c[14] * x * y;
    r += kShC[5] * c[15] * y * z; g += kShC[5] * c[16] * y * z; b += kShC[5] * c[17] * y * z;
    const float t6 = 3.0f * zz - 1.0f;
    r += kShC[6] * c[18] * t6; g += kShC[6] * c[19] * t6; b += kShC[6] * c[20] * t6;
    r += kShC[7] * c[21] * x * z; g += kShC[7] * c[22] * x * z; b += kShC[7] * c[23] * x * z;
    const float t8 = xx - yy;
    r += kShC[8] * c[24] * t8; g += kShC[8] * c[25] * t8; b += kShC[8] * c[26] * t8;
  }
  if (DEG >= 3) {
    const float x = d.x, y = d.y, z = d.z;
    const float xx = x * x, yy = y * y, zz = z * z;
    const float f[7] = {y * (3.0f * xx - yy), x * y * z,           y * (5.0f * zz - 1.0f), z * (5.0f * zz - 3.0f),
                        x * (5.0f * zz - 1.0f), z * (xx - yy), x * (xx - 3.0f * yy)};
#pragma unroll
    for (int k = 9; k <= 15; ++k) {
      r += kShC[k] * c[3 * k] * f[k - 9];
      g += kShC[k] * c[3 * k + 1] * f[k - 9];
      b += kShC[k] * c[3 * k + 2] * f[k - 9];
    }
  }
  return V3{0.5f + r, 0.5f + g, 0.5f + b};
}

__device__ __forceinline__ V3 eval_sh_dyn(int deg, const float* __restrict__ c, V3 d) {
  switch (deg) {
    case 0: return eval_sh_row<0>(c, d);
    case 1: return eval_sh_row<1>(c, d);
    case 2: return eval_sh_row<2>(c, d);
    default: return eval_sh_row<3>(c, d);
  }
}

// d colour_c / d coef(k, c) = C_k * factor_k(d) (the adjoint's SH basis).
__device__ __forceinline__ void sh_basis(int K, V3 d, float out[16]) {
  const float x = d.x, y = d.y, z = d.z;
  const float xx = x * x, yy = y * y, zz = z * z;
  const float f[16] = {1.0f,
                       y,
                       z,
                       x,
                       x * y,
                       y * z,
                       3.0f * zz - 1.0f,
                       x * z,
                       xx - yy,
                       y * (3.0f * xx - yy),
                       x * y * z,
                       y * (5.0f * zz - 1.0f),
                       z * (5.0f * zz - 3.0f),
                       x * (5.0f * zz - 1.0f),
                       z * (xx - yy),
                       x * (xx - 3.0f * yy)};
#pragma unroll
  for (int k = 0; k < 16; ++k) out[k] = k < K ? kShC[k] * f[k] : 0.0f;
}

// Pixel-centre camera ray: radfoam.wgsl:45-52 with qrot (common.wgsl:19-23),
// normalised by division.  Shared by bv_camera_rays and the camera kernels.
__device__ __forceinline__ void camera_ray(const RaySource& rs, uint32_t x, uint32_t y, V3* o, V3* d) {
  const float px = ((float)x + 0.5f) / (float)rs.width;
  const float py = ((float)y + 0.5f) / (float)rs.height;
  const float nx = px * 2.0f - 1.0f, ny = py * 2.0f - 1.0f;
  const V3 v = mk((nx - rs.principal[0]) * rs.tan_half[0], (ny - rs.principal[1]) * rs.tan_half[1], 1.0f);
  const V3 q = mk(rs.quat[0], rs.quat[1], rs.quat[2]);
  const V3 t = scl(2.0f, cross(q, v));
  const V3 r = add(add(v, scl(rs.quat[3], t)), cross(q, t));
  const float len = sqrtf(dot(r, r));
  *o = mk(rs.cam_pos[0], rs.cam_pos[1], rs.cam_pos[2]);
  *d = mk(r.x / len, r.y / len, r.z / len);
}

// Ray i of a source.  Camera sources map a linear index to 8x8 pixel tiles so a
// 64-lane wave covers one tile (neighbouring rays walk neighbouring cells).
// A per-ray start cell outside the scene (BV_NO_CELL included) makes the ray
// empty: its direction becomes zero, which every walk treats as a degenerate
// ray (no cell is read, zero colour, zero steps, status 0, no gradient).
__device__ __forceinline__ bool fetch_ray(const RaySource& rs, int64_t i, V3* o, V3* d, uint32_t* start,
                                          int64_t* out_index, bool /*writer*/ = true) {
  if (rs.origins != nullptr) {
    *o = mk(rs.origins[3 * i], rs.origins[3 * i + 1], rs.origins[3 * i + 2]);
    *d = mk(rs.dirs[3 * i], rs.dirs[3 * i + 1], rs.dirs[3 * i + 2]);
    *start = rs.start_cells ? rs.start_cells[i] : rs.start_cell;
    if (rs.n_cells != 0u && *start >= rs.n_cells) *d = mk(0.f, 0.f, 0.f);
    *out_index = i;
    return true;
  }
  const uint32_t tiles_x = (rs.width + 7u) / 8u;
  const uint64_t tile = (uint64_t)i >> 6;
  const uint32_t lane = (uint32_t)i & 63u;
  const uint32_t x = (uint32_t)(tile % tiles_x) * 8u + (lane & 7u);
  const uint32_t y = (uint32_t)(tile / tiles_x) * 8u + (lane >> 3);
  if (x >= rs.width || y >= rs.height) return false;
  camera_ray(rs, x, y, o, d);
  *start = rs.start_cell;
  *out_index = (int64_t)y * rs.width + x;
  return true;
}

// Number of linear ray slots a source spans (camera sources are tile-padded).
__host__ __device__ inline int64_t ray_slots(const RaySource& rs) {
  if (rs.origins != nullptr) return rs.count;
  const int64_t tx = (rs.width + 7) / 8, ty = (rs.height + 7) / 8;
  return tx * ty * 64;
}

// trace.rs:1704-1706: divide by the length; false for a degenerate direction.
__device__ __forceinline__ bool normalise(V3* d) {
  const float len = sqrtf(dot(*d, *d));
  if (!(len > 0.0f) || !isfinite(len)) return false;
  *d = mk(d->x / len, d->y / len, d->z / len);
  return true;
}

// ---------------------------------------------------------------- batched recorder rays
// One camera slice of a batched path-record dispatch (RecordPathsArgs,
// gpu/path_record.rs:142-161) in device form.  The half-angle tangents are
// computed on the host, once per slice.
struct ViewSlice {
  float pos[3];
  float quat[4];
  float tan_half[2];
  float principal[2];
  uint32_t width, height;
  uint32_t start;               // start_point
  uint32_t pixel_offset;        // first output row (and first pixel_indices entry) of the slice
  uint32_t image_pixel_offset;  // added to every pixel index
  uint32_t view;                // index of the slice in the caller's list (view_index output)
  uint32_t pad;
  int64_t slot_begin;           // first linear slot of the slice within its launch
};
constexpr int kViewsPerLaunch = 24;

// Ray source of one recorder launch: up to kViewsPerLaunch slices, laid end to
// end in slot order.  Slot i of slice k writes output row
// pixel_offset_k + (i - slot_begin_k).  The table travels as a kernel argument
// (the kernarg segment), so no device allocation or upload is needed.
struct ViewRays {
  const uint32_t* __restrict__ pixel_indices;  // [rows] indexed by output row
  float* ray_origins;   // [rows][3] or null: per-row origin walked
  float* ray_dirs;      // [rows][3] or null: per-row unit direction walked
  uint32_t* view_index; // [rows] or null: the row's slice (ViewSlice::view)
  int64_t count;        // slots of this launch
  uint32_t n;           // slices in use
  uint32_t pad;
  ViewSlice v[kViewsPerLaunch];
};
static_assert(sizeof(ViewSlice) == 80 && sizeof(ViewRays) <= 2048, "recorder view table must stay a small kernel argument");

// radfoam_record_paths.wgsl:133-137: the recorder's own quaternion rotation,
// 2 (u.v) u + (s^2 - u.u) v + 2 s (u x v), evaluated left to right.
__device__ __forceinline__ V3 qrot_record(const float q[4], V3 v) {
  const V3 u = mk(q[0], q[1], q[2]);
  const float s = q[3];
  const V3 a = scl(2.0f * dot(u, v), u);
  const V3 b = scl(s * s - dot(u, u), v);
  const V3 c = scl(2.0f * s, cross(u, v));
  return add(add(a, b), c);
}

__host__ __device__ inline int64_t ray_slots(const ViewRays& rs) { return rs.count; }

// Slice holding slot i: slices are in ascending slot order, so it is the
// number of later slices whose first slot is <= i (constant indices only).
__device__ __forceinline__ uint32_t view_of_slot(const ViewRays& rs, int64_t i) {
  uint32_t k = 0;
#pragma unroll
  for (int j = 1; j < kViewsPerLaunch; ++j) k += ((uint32_t)j < rs.n && i >= rs.v[j].slot_begin) ? 1u : 0u;
  return k;
}

// record_paths / record_powerfoam_splats ray set-up (radfoam_record_paths.wgsl
// :1227-1230, :1113-1116): pixel = pixel_indices[row] + image_pixel_offset
// (u32 arithmetic), origin = camera position, direction =
// qrot(orientation, ((ndc - principal) * tan_half, 1)) from ray_dir_for_pixel
// (:139-153), returned unnormalised: every walk divides by its length once,
// which is the shader's normalize.  With `writer` set, the row's origin, unit
// direction and slice are stored to the optional per-row outputs.
__device__ __forceinline__ bool fetch_ray(const ViewRays& rs, int64_t i, V3* o, V3* d, uint32_t* start,
                                          int64_t* out_index, bool writer = true) {
  const uint32_t k = view_of_slot(rs, i);
  const ViewSlice& v = rs.v[k];
  const int64_t row = (int64_t)v.pixel_offset + (i - v.slot_begin);
  const uint32_t pidx = rs.pixel_indices[row] + v.image_pixel_offset;
  const uint32_t ix = pidx % v.width, iy = pidx / v.width;
  const float px = ((float)ix + 0.5f) / (float)v.width;
  const float py = ((float)iy + 0.5f) / (float)v.height;
  const float nx = px * 2.0f - 1.0f, ny = py * 2.0f - 1.0f;
  const V3 local = mk((nx - v.principal[0]) * v.tan_half[0], (ny - v.principal[1]) * v.tan_half[1], 1.0f);
  *o = mk(v.pos[0], v.pos[1], v.pos[2]);
  *d = qrot_record(v.quat, local);
  *start = v.start;
  *out_index = row;
  if (writer) {
    if (rs.ray_origins) {
      rs.ray_origins[3 * row] = o->x;
      rs.ray_origins[3 * row + 1] = o->y;
      rs.ray_origins[3 * row + 2] = o->z;
    }
    if (rs.ray_dirs) {
      V3 u = *d;
      normalise(&u);
      rs.ray_dirs[3 * row] = u.x;
      rs.ray_dirs[3 * row + 1] = u.y;
      rs.ray_dirs[3 * row + 2] = u.z;
    }
    if (rs.view_index) rs.view_index[row] = v.view;
  }
  return true;
}

// support_interval, trace.rs:209-238.
__device__ __forceinline__ bool support_interval(bool bounded, V3 ro, V3 rd, V3 c, float radius, float t0,
                                                 float t1, float* s0, float* s1) {
  if (!bounded) {
    *s0 = t0;
    *s1 = t1;
    return t1 > t0;
  }
  const V3 oc = sub(ro, c);
  const float a = dot(rd, rd);
  if (!isfinite(a) || a <= 0.0f) return false;
  const float b = dot(oc, rd);
  const float cc = dot(oc, oc) - radius * radius;
  const float disc = b * b - a * cc;
  if (disc <= 0.0f) return false;
  const float root = sqrtf(disc);
  const float nt = (-b - root) / a;
  const float ft = (-b + root) / a;
  const float cs = t0 > nt ? t0 : nt;
  const float ce = t1 < ft ? t1 : ft;
  *s0 = cs;
  *s1 = ce;
  return ce > cs;
}

// One plane test (find_exit_face body, trace.rs:186-202): offset o = p_j - p_i,
// face origin p_i + shift * o, crossing t = ((fo - ro) . o) / (o . d) if o . d > 0.
__device__ __forceinline__ void plane_test(V3 cp, V3 o, float shift, V3 ro, V3 d, float t0, uint32_t nb,
                                           float* best, uint32_t* best_nb) {
  const V3 fo = add(cp, scl(shift, o));
  const float dp = dot(o, d);
  if (dp > 0.0f) {
    const float t = dot(sub(fo, ro), o) / dp;
    if (isfinite(t) && t > t0 && t < *best) {
      *best = t;
      *best_nb = nb;
    }
  }
}


// ---------------------------------------------------------------- sphere BVH
// Linear BVH over the support spheres of a power-cell scene (built by the
// splat tracer).  Inner node: the boxes of both children and their references;
// a reference with kLeaf set is a leaf position in Morton order.
constexpr uint32_t kLeaf = 0x80000000u;
struct __align__(16) BvhNode {
  float4 a;  // lmin.x lmin.y lmin.z lmax.x
  float4 b;  // lmax.y lmax.z rmin.x rmin.y
  float4 c;  // rmin.z rmax.x rmax.y rmax.z
  uint4 d;   // left right parent -
};

struct BvhView {
  const BvhNode* __restrict__ nodes;  // [max(n - 1, 1)]
  const CellRec* __restrict__ leaf;   // [n] cell records in Morton order
  const uint32_t* __restrict__ cell;  // [n] site index of each leaf
  uint32_t stack_rows;                // traversal stack entries per lane: tree depth + 1
};

// ---------------------------------------------------------------- relighting
// shade_point_light (relight.rs:295-324) with PointLight::sample (:203-233)
// and point_light_specular (:249-293).
__device__ __forceinline__ bool try_normalize(V3* v) {
  const float len = sqrtf(dot(*v, *v));
  if (!(len > 0.0f) || !isfinite(len)) return false;
  *v = scl(1.0f / len, *v);
  return true;
}

__device__ __forceinline__ V3 shade_point_light(const bv_point_light& L, V3 point, V3 normal, V3 view, float4 mat,
                                                const float f0[3]) {
  const V3 zero = mk(0.f, 0.f, 0.f);
  const V3 from = sub(point, mk(L.position[0], L.position[1], L.position[2]));
  const float dsq = dot(from, from);
  if (!isfinite(dsq) || dsq <= 1.1920929e-7f || !(L.exponent >= 0.0f) || !isfinite(L.exponent)) return zero;
  if (!(L.intensity[0] >= 0.f) || !(L.intensity[1] >= 0.f) || !(L.intensity[2] >= 0.f) || !isfinite(L.intensity[0]) ||
      !isfinite(L.intensity[1]) || !isfinite(L.intensity[2]))
    return zero;
  const float inv_d = 1.0f / sqrtf(dsq);
  const V3 away = mk(from.x * inv_d, from.y * inv_d, from.z * inv_d);
  float angular = 1.0f;
  if (L.exponent != 0.0f) {
    V3 dl = mk(L.direction[0], L.direction[1], L.direction[2]);
    if (!try_normalize(&dl)) return zero;
    const float cdot = dot(dl, away);
    angular = powf(cdot > 0.f ? cdot : 0.f, L.exponent);
  }
  const V3 towards = mk(-away.x, -away.y, -away.z);
  const V3 radiance = mk(L.intensity[0] * angular / dsq, L.intensity[1] * angular / dsq, L.intensity[2] * angular / dsq);
  V3 n = normal;
  const bool n_ok = try_normalize(&n);
  float cosine = 0.f;
  if (n_ok) {
    cosine = dot(n, towards);
    cosine = cosine > 0.f ? cosine : 0.f;
  }
  V3 out = mk(mat.x * radiance.x * cosine, mat.y * radiance.y * cosine, mat.z * radiance.z * cosine);
  if (f0[0] == 0.f && f0[1] == 0.f && f0[2] == 0.f) return out;
  V3 v = view;
  if (!n_ok || !try_normalize(&v)) return out;
  float n_dot_l = dot(n, towards);
  n_dot_l = n_dot_l > 0.f ? n_dot_l : 0.f;
  float n_dot_v = dot(n, v);
  n_dot_v = n_dot_v > 0.f ? n_dot_v : 0.f;
  V3 h = add(towards, v);
  if (!try_normalize(&h) || n_dot_l == 0.f || n_dot_v == 0.f) return out;
  float n_dot_h = dot(n, h);
  n_dot_h = n_dot_h > 0.f ? n_dot_h : 0.f;
  float v_dot_h = dot(v, h);
  v_dot_h = v_dot_h < 0.f ? 0.f : (v_dot_h > 1.f ? 1.f : v_dot_h);
  const float r = mat.w < 0.045f ? 0.045f : (mat.w > 1.f ? 1.f : mat.w);
  const float alpha = r * r;
  const float a2 = alpha * alpha;
  const float den = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
  const float pi = 3.14159265358979323846f;
  float dd = pi * den * den;
  dd = dd > 1.17549435e-38f ? dd : 1.17549435e-38f;
  const float distribution = a2 / dd;
  const float sl = 2.0f * n_dot_l / (n_dot_l + sqrtf(a2 + (1.0f - a2) * n_dot_l * n_dot_l));
  const float sv = 2.0f * n_dot_v / (n_dot_v + sqrtf(a2 + (1.0f - a2) * n_dot_v * n_dot_v));
  const float g1 = 1.0f - v_dot_h;
  const float grazing = g1 * g1 * g1 * g1 * g1;
  const float scale = pi * distribution * (sl * sv) / (4.0f * n_dot_v);
  out.x += radiance.x * scale * (f0[0] + (1.0f - f0[0]) * grazing);
  out.y += radiance.y * scale * (f0[1] + (1.0f - f0[1]) * grazing);
  out.z += radiance.z * scale * (f0[2] + (1.0f - f0[2]) * grazing);
  return out;
}


// One admitted view request of the serving loop, as the tile-queue kernel sees
// it: the scene (and its sphere BVH for splat requests), the camera (with its
// start cell), the settings, the lights and the output image in the server's
// arena.
struct ServeReq {
  SceneView s;
  BvhView bvh;
  RaySource rs;
  Settings st;
  bv_lighting lit;
  float* out;
};

// ---------------------------------------------------------------- serving queue
// The serving kernel is persistent: its blocks take 8x8 ray tiles from a ring
// that the server's scheduler thread keeps filling while the kernel runs.  The
// ring, its published count, the request table and the per-entry completion
// words live in pinned host memory (read and written by the kernel through
// PCIe, so no copy kernel has to find a free CU while this one runs); block 0
// mirrors the host's count and stop flag into device memory, so blocks that
// wait for work poll HBM, not PCIe.
struct ServeCtl {
  uint64_t published;  // ring entries the host has written
  uint32_t stop;       // 1: leave once every published entry is taken
  uint32_t heartbeat;  // advanced by the scheduler thread while it runs
};

struct ServeQueue {
  const ServeCtl* ctl;      // pinned host memory
  const uint32_t* ring;     // pinned host memory [mask + 1]: (request slot << 20) | tile
  const ServeReq* reqs;     // pinned host memory [slots]
  uint32_t* done;           // pinned host memory [mask + 1]: entry index + 1 once its tile is written
  uint64_t* claimed;        // device: entries taken
  uint64_t* published_dev;  // device: the mirror of ctl->published
  uint32_t* stop_dev;       // device: the mirror of ctl->stop
  uint32_t mask;
  uint64_t dead_ticks;      // wall-clock ticks without a heartbeat after which the kernel leaves
};

constexpr uint32_t kServeNoTile = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t sys_load(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t sys_load(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint64_t dev_load(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t dev_load(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Block 0's first lane: copy the host's published count, then its stop flag,
// into device memory until the host stops the kernel or its heartbeat has not
// moved for q.dead_ticks (the process is gone or wedged).  Every other block
// leaves once it sees the stop and has no published entry left to take, so the
// whole grid drains either way.
__device__ __forceinline__ void serve_mirror(const ServeQueue& q) {
  if (threadIdx.x != 0) return;
  uint32_t beat = sys_load(&q.ctl->heartbeat);
  uint64_t beat_at = wall_clock64();
  for (;;) {
    const uint32_t stop = sys_load(&q.ctl->stop);
    // the count after the stop flag: the host publishes nothing after stopping
    __hip_atomic_store(q.published_dev, sys_load(&q.ctl->published), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t b = sys_load(&q.ctl->heartbeat);
    const uint64_t now = wall_clock64();
    if (b != beat) {
      beat = b;
      beat_at = now;
    }
    if (stop != 0u || now - beat_at > q.dead_ticks) {
      __hip_atomic_store(q.stop_dev, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(32);
  }
}

// The block's next ring entry: (entry index, tile code), or kServeNoTile once
// the kernel is to leave.  Lane 0 of the block claims and waits; the caller's
// block reads the result from `slot_*` after a barrier.
__device__ __forceinline__ void serve_claim(const ServeQueue& q, uint64_t* slot_entry, uint32_t* slot_code) {
  if (threadIdx.x == 0) {
    const uint64_t c = atomicAdd(reinterpret_cast<unsigned long long*>(q.claimed), 1ull);
    uint32_t code = kServeNoTile;
    uint32_t nap = 1;
    for (;;) {
      if (c < dev_load(q.published_dev)) {
        code = sys_load(q.ring + (c & q.mask));
        break;
      }
      if (dev_load(q.stop_dev) != 0u && c >= dev_load(q.published_dev)) break;
      // a doubling nap of up to about 4 us between polls
      for (uint32_t k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(8);
      nap = nap < 16u ? 2u * nap : 16u;
    }
    *slot_entry = c;
    *slot_code = code;
  }
  __syncthreads();
}

// The request's descriptor into LDS (word by word, bypassing the caches: the
// host rewrites a slot's descriptor between requests).
__device__ __forceinline__ void serve_load_req(const ServeQueue& q, uint32_t slot, ServeReq* lreq) {
  static_assert(sizeof(ServeReq) % 4 == 0, "descriptor copied in 32-bit words");
  const uint32_t* src = reinterpret_cast<const uint32_t*>(q.reqs + slot);
  uint32_t* dst = reinterpret_cast<uint32_t*>(lreq);
  for (uint32_t i = threadIdx.x; i < sizeof(ServeReq) / 4; i += blockDim.x) dst[i] = sys_load(src + i);
  __syncthreads();
}

// A descriptor field read from LDS into scalar registers (every lane holds
// the same words), so the tile's walk keeps its vector registers.
template <class T>
__device__ __forceinline__ T uniform_copy(const T& lds) {
  static_assert(sizeof(T) % 4 == 0, "copied in 32-bit words");
  T out;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&lds);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (uint32_t i = 0; i < sizeof(T) / 4; ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

// After the tile's pixels are stored: make them visible to the host, then mark
// the entry done.
__device__ __forceinline__ void serve_done(const ServeQueue& q, uint64_t entry) {
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(q.done + (entry & q.mask), (uint32_t)(entry + 1), __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

// Front-to-back accumulation state shared by every walk kernel.
struct WalkAcc {
  float r, g, b;
  float trans;
  float depth_mode, peak;
  uint32_t peak_point;
  uint32_t bytes, segments;  // instrumentation (algorithmic bytes, integrated segments)
};

// Algorithmic bytes of one visited cell: CSR bounds, neighbour ids and
// positions (and radii for power cells), the density, and the SH row when the
// cell is integrated.
__device__ __forceinline__ uint32_t step_bytes(uint32_t deg, bool bounded) {
  return 8u + deg * (bounded ? 20u : 16u) + 4u;
}

template <int EVAL>
__device__ __forceinline__ void integrate(const Settings& st, float sd, float seg0, float seg1, uint32_t cur, V3 col,
                                          WalkAcc* a) {
  const float dt = seg1 - seg0;
  const float alpha = 1.0f - expf(-sd * dt);
  const float w = a->trans * alpha;
  if (EVAL == BV_EVAL_CONSTANT) {
    a->r += w * st.const_rgb[0];
    a->g += w * st.const_rgb[1];
    a->b += w * st.const_rgb[2];
  } else if (EVAL == BV_EVAL_SH) {
    a->r += w * col.x;
    a->g += w * col.y;
    a->b += w * col.z;
  }
  if (w > a->peak) {
    a->peak = w;
    a->depth_mode = 0.5f * (seg0 + seg1);
    a->peak_point = cur;
  }
  a->trans *= 1.0f - alpha;
  a->segments += 1;
}

__device__ __forceinline__ V3 clamp0(V3 c) {
  return mk(c.x > 0.0f ? c.x : 0.0f, c.y > 0.0f ? c.y : 0.0f, c.z > 0.0f ? c.z : 0.0f);
}

__device__ __forceinline__ void store_out(const Outputs& out, int64_t idx, const WalkAcc& a, uint32_t steps,
                                          uint32_t last, float t_end) {
  if (out.rgba) reinterpret_cast<float4*>(out.rgba)[idx] = make_float4(a.r, a.g, a.b, 1.0f - a.trans);
  if (out.steps) out.steps[idx] = steps;
  if (out.last_point) out.last_point[idx] = last;
  if (out.t_end) out.t_end[idx] = t_end;
  if (out.depth_mode) out.depth_mode[idx] = a.depth_mode;
  if (out.peak_weight) out.peak_weight[idx] = a.peak;
  if (out.peak_point) out.peak_point[idx] = a.peak_point;
  if (out.stats) reinterpret_cast<uint2*>(out.stats)[idx] = make_uint2(a.bytes + 16u, a.segments);
}

}  // namespace bvk
