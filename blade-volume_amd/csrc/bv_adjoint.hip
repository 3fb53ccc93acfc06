// Adjoint of the forward walk (SH colour, Voronoi cells): one replay per ray,
// nothing stored between forward and backward.
//
// The trainer's differentiable forward (blade-volume-train/src/diff_render.rs:
// 1306-1572, 2162-2355) integrates recorded (cell, next, dt) rows with
// tau = sigma dt, T = exp(-exclusive_cumsum(tau)), alpha = 1 - exp(-tau),
// w = T alpha, colour = relu(0.5 + SH), pixel = sum w c, with dt the difference
// of consecutive bisector-plane crossings.  Its gradient, which meganeura forms
// with embedding scatter-adds, is here one kernel.  Given the forward's final
// colour C and transmittance T_N, a single forward-order replay yields, for the
// integrated segment k,
//   dL/dtau_k = sum_c g_c (T_{k+1} c_k - (C - C_{<=k})) + g_a T_N
//   dL/dsigma_k = dL/dtau_k dt_k,      dL/ddt_k = dL/dtau_k sigma_k
//   dL/dcoef(k, j, c) = g_c w_k [c unclamped > 0] C_j Y_j(d)
// and the positions get dL/ddt_k times the closed-form plane-crossing
// Jacobian of trace.rs:1077-1097 at the exit face (current, next) minus the one
// at the entry face (previous, current).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

#include "bv_host.h"

namespace {

using namespace bvk;

struct Grads {
  float* d_density;
  float* d_sh;
  float* d_xyz;
};

// Derivative of a bisector crossing t with respect to the two sites
// (face_intersection_jacobians, trace.rs:1077-1097, radius terms dropped).
__device__ __forceinline__ void face_jac(V3 ro, V3 rd, V3 cur, V3 adj, float t, V3* d_cur, V3* d_adj) {
  const V3 nrm = sub(adj, cur);
  const float den = dot(rd, nrm);
  if (!isfinite(den) || fabsf(den) <= 1e-20f) {
    *d_cur = mk(0.f, 0.f, 0.f);
    *d_adj = mk(0.f, 0.f, 0.f);
    return;
  }
  const V3 a = add(sub(ro, cur), scl(t, rd));
  const V3 b = sub(sub(adj, ro), scl(t, rd));
  *d_cur = mk(a.x / den, a.y / den, a.z / den);
  *d_adj = mk(b.x / den, b.y / den, b.z / den);
}

__device__ __forceinline__ void atomic_add3(float* p, V3 v) {
  atomicAdd(p, v.x);
  atomicAdd(p + 1, v.y);
  atomicAdd(p + 2, v.z);
}

struct SegmentGrad {
  float d_sigma, d_dt;
  float gc[3];
  bool any_colour;
};

// Replay bookkeeping shared by both kernels.
struct AdjState {
  float Cr, Cg, Cb, Tfin;
  float g[4];
  float trans, pr, pg, pb;
};

__device__ __forceinline__ SegmentGrad segment_grad(AdjState* st, float sd, float dt, V3 cu) {
  const float alpha = 1.0f - expf(-sd * dt);
  const float w = st->trans * alpha;
  const float cr = cu.x > 0.f ? cu.x : 0.f, cg = cu.y > 0.f ? cu.y : 0.f, cb = cu.z > 0.f ? cu.z : 0.f;
  st->pr += w * cr;
  st->pg += w * cg;
  st->pb += w * cb;
  const float tn = st->trans * (1.0f - alpha);
  const float gtau = st->g[0] * (tn * cr - (st->Cr - st->pr)) + st->g[1] * (tn * cg - (st->Cg - st->pg)) +
                     st->g[2] * (tn * cb - (st->Cb - st->pb)) + st->g[3] * st->Tfin;
  SegmentGrad r;
  r.d_sigma = gtau * dt;
  r.d_dt = gtau * sd;
  r.gc[0] = cu.x > 0.f ? st->g[0] * w : 0.f;
  r.gc[1] = cu.y > 0.f ? st->g[1] * w : 0.f;
  r.gc[2] = cu.z > 0.f ? st->g[2] * w : 0.f;
  r.any_colour = r.gc[0] != 0.f || r.gc[1] != 0.f || r.gc[2] != 0.f;
  st->trans = tn;
  return r;
}

__device__ __forceinline__ void scatter_sh(float* dsh_row, int K, const float basis[16], const float gc[3]) {
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (gc[c] != 0.f) atomicAdd(dsh_row + 3 * k + c, gc[c] * basis[k]);
}

__device__ __forceinline__ bool adj_begin(const RaySource& rs, int64_t i, const float* rgba_fwd, const float* d_rgba,
                                          V3* ro, V3* dir, uint32_t* cur, AdjState* st) {
  int64_t oi;
  if (!fetch_ray(rs, i, ro, dir, cur, &oi)) return false;
  if (!normalise(dir)) return false;
  const float4 C = reinterpret_cast<const float4*>(rgba_fwd)[oi];
  const float4 g = reinterpret_cast<const float4*>(d_rgba)[oi];
  st->Cr = C.x; st->Cg = C.y; st->Cb = C.z; st->Tfin = 1.0f - C.w;
  st->g[0] = g.x; st->g[1] = g.y; st->g[2] = g.z; st->g[3] = g.w;
  st->trans = 1.0f; st->pr = 0.f; st->pg = 0.f; st->pb = 0.f;
  return true;
}

// ---------------------------------------------------------------- plain port
// One thread per ray on the reference layout; every gradient element is its
// own global atomic (the scatter-add of the reference's embedding backward).
__global__ void __launch_bounds__(64) k_adjoint_plain(SceneView s, RaySource rs, Settings st, const float* rgba_fwd,
                                                      const float* d_rgba, Grads gr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rs.count) return;
  V3 ro, dir;
  uint32_t cur;
  AdjState a;
  if (!adj_begin(rs, i, rgba_fwd, d_rgba, &ro, &dir, &cur, &a)) return;
  float basis[16];
  sh_basis(s.K, dir, basis);
  float4 p = s.pts[cur];
  V3 cp = mk(p.x, p.y, p.z), pp = cp;
  uint32_t prev = cur;
  float t0 = 0.0f;
  uint32_t steps = 0;
  while (steps < st.max_steps) {
    steps += 1;
    if (a.trans <= st.weight_threshold || t0 > st.depth) break;
    const uint32_t b = s.offsets[cur], e = s.offsets[cur + 1];
    float t1 = st.depth;
    uint32_t nb_best = BV_NO_CELL;
    for (uint32_t j = b; j < e; ++j) {
      const uint32_t nb = s.neighbors[j];
      const float4 q = s.pts[nb];
      const V3 off = sub(mk(q.x, q.y, q.z), cp);
      float dsq = dot(off, off);
      dsq = dsq > 1e-20f ? dsq : 1e-20f;
      const float shift = 0.5f + 0.5f * (0.0f - q.w * q.w) / dsq;
      plane_test(cp, off, shift, ro, dir, t0, nb, &t1, &nb_best);
    }
    const uint32_t nxt = nb_best == BV_NO_CELL ? cur : nb_best;
    float4 pn = s.pts[nxt];
    if (t1 > t0) {
      const float* row = s.attrs + (int64_t)cur * s.attr_dim;
      const float sd = row[3 * s.K];
      if (sd > 1e-6f) {
        const float dt = t1 - t0;
        const V3 cu = eval_sh_dyn(s.sh_degree, row, dir);
        const SegmentGrad sg = segment_grad(&a, sd, dt, cu);
        atomicAdd(gr.d_density + cur, sg.d_sigma);
        if (sg.any_colour) scatter_sh(gr.d_sh + (int64_t)cur * 3 * s.K, s.K, basis, sg.gc);
        if (nb_best != BV_NO_CELL) {
          V3 dc, dn;
          face_jac(ro, dir, cp, mk(pn.x, pn.y, pn.z), t1, &dc, &dn);
          atomic_add3(gr.d_xyz + 3 * (int64_t)cur, scl(sg.d_dt, dc));
          atomic_add3(gr.d_xyz + 3 * (int64_t)nxt, scl(sg.d_dt, dn));
        }
        if (prev != cur) {
          V3 dp, dc;
          face_jac(ro, dir, pp, cp, t0, &dp, &dc);
          atomic_add3(gr.d_xyz + 3 * (int64_t)prev, scl(-sg.d_dt, dp));
          atomic_add3(gr.d_xyz + 3 * (int64_t)cur, scl(-sg.d_dt, dc));
        }
      }
    }
    if (nb_best == BV_NO_CELL) break;
    t0 = t0 > t1 ? t0 : t1;
    prev = cur;
    pp = cp;
    cur = nxt;
    cp = mk(pn.x, pn.y, pn.z);
  }
}

__global__ void k_l2_loss_grad(const float4* __restrict__ rgba, const float* __restrict__ target, int64_t n,
                               float4* __restrict__ d_rgba, float* loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    const float4 c = rgba[i];
    const float er = c.x - target[3 * i], eg = c.y - target[3 * i + 1], eb = c.z - target[3 * i + 2];
    d_rgba[i] = make_float4(2.f * er, 2.f * eg, 2.f * eb, 0.f);
    l = (er * er + eg * eg) + eb * eb;
  }
  if (loss) {
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
  }
}

}  // namespace

extern "C" {

int bv_adjoint(bv_scene* s, const bv_rays* rays, const bv_trace_settings* settings, const float* rgba_forward,
               const float* d_rgba, const bv_grads* grads, int kernel, void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(settings && grads, BV_ERR_INVALID, "null settings/grads");
  BV_REQUIRE(settings->eval_mode == BV_EVAL_SH, BV_ERR_UNSUPPORTED, "the adjoint differentiates SH colour only");
  BV_REQUIRE(!s->bounded, BV_ERR_UNSUPPORTED, "the adjoint supports Voronoi (radius-free) cells only");
  BV_REQUIRE(kernel == BV_KERNEL_PLAIN || kernel == BV_KERNEL_FAST, BV_ERR_INVALID, "bad kernel id");
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rgba_forward && d_rgba, BV_ERR_INVALID, "null forward colour / loss gradient");
  BV_REQUIRE(grads->d_density && grads->d_sh && grads->d_xyz, BV_ERR_INVALID, "null gradient table");
  const bvk::RaySource rs = bvh::walk_rays(s, rays);
  const bvk::Settings st = bvh::settings_of(settings);
  const Grads g{grads->d_density, grads->d_sh, grads->d_xyz};
  const bvk::SceneView v = s->view();
  hipStream_t sm = (hipStream_t)stream;
  if (kernel == BV_KERNEL_PLAIN) {
    const int bs = 64;
    hipLaunchKernelGGL(k_adjoint_plain, dim3((unsigned)((rays->count + bs - 1) / bs)), dim3(bs), 0, sm, v, rs, st,
                       rgba_forward, d_rgba, g);
  } else {
    return bvh::launch_adjoint_coop(s, rs, st, rgba_forward, d_rgba,
                                    bvh::GradPtrs{grads->d_density, grads->d_sh, grads->d_xyz}, 0, sm);
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int bv_train_step_l2(bv_scene* s, const bv_rays* rays, const bv_trace_settings* settings, const float* target_rgb,
                     float* rgba_out, const bv_grads* grads, float* loss, int kernel, void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(settings && grads, BV_ERR_INVALID, "null settings/grads");
  BV_REQUIRE(settings->eval_mode == BV_EVAL_SH, BV_ERR_UNSUPPORTED, "training differentiates SH colour only");
  BV_REQUIRE(!s->bounded, BV_ERR_UNSUPPORTED, "the adjoint supports Voronoi (radius-free) cells only");
  BV_REQUIRE(kernel == BV_KERNEL_PLAIN || kernel == BV_KERNEL_FAST, BV_ERR_INVALID, "bad kernel id");
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rays->count < (int64_t(1) << 31), BV_ERR_INVALID, "a training call takes fewer than 2^31 rays");
  BV_REQUIRE(target_rgb != nullptr, BV_ERR_INVALID, "null target colours");
  BV_REQUIRE(grads->d_density && grads->d_sh && grads->d_xyz, BV_ERR_INVALID, "null gradient table");
  hipStream_t sm = (hipStream_t)stream;
  const bvk::RaySource rs = bvh::walk_rays(s, rays);
  const bvk::Settings st = bvh::settings_of(settings);
  const bvh::GradPtrs g{grads->d_density, grads->d_sh, grads->d_xyz};
  if (kernel == BV_KERNEL_FAST) return bvh::launch_train_l2_coop(s, rs, st, target_rgb, rgba_out, loss, g, 0, sm);
  // Plain pipeline: the reference's shape (walk, loss, backward) as three launches.
  void* scratch = nullptr;
  rc = bvh::stream_buffer(s, s->scratch_bufs, sm, (size_t)rays->count * 32, &scratch);
  if (rc != BV_OK) return rc;
  float* rgba = rgba_out ? rgba_out : reinterpret_cast<float*>(scratch);
  float* d_rgba = reinterpret_cast<float*>(scratch) + 4 * rays->count;
  bv_trace_outputs out{};
  out.rgba = rgba;
  rc = bv_trace_rays(s, rays, settings, &out, BV_KERNEL_PLAIN, stream);
  if (rc != BV_OK) return rc;
  rc = bv_l2_loss_grad(rgba, target_rgb, rays->count, d_rgba, loss, stream);
  if (rc != BV_OK) return rc;
  return bv_adjoint(s, rays, settings, rgba, d_rgba, grads, BV_KERNEL_PLAIN, stream);
}

int bv_train_step_l2_traced(bv_scene* s, const bv_rays* rays, const bv_trace_settings* settings,
                            const float* target_rgb, float* rgba_out, const bv_grads* grads, float* loss,
                            uint32_t* path_cells, uint32_t path_len, void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(settings && grads, BV_ERR_INVALID, "null settings/grads");
  BV_REQUIRE(settings->eval_mode == BV_EVAL_SH, BV_ERR_UNSUPPORTED, "training differentiates SH colour only");
  BV_REQUIRE(!s->bounded, BV_ERR_UNSUPPORTED, "the adjoint supports Voronoi (radius-free) cells only");
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rays->count < (int64_t(1) << 31), BV_ERR_INVALID, "a training call takes fewer than 2^31 rays");
  BV_REQUIRE(target_rgb != nullptr, BV_ERR_INVALID, "null target colours");
  BV_REQUIRE(grads->d_density && grads->d_sh && grads->d_xyz, BV_ERR_INVALID, "null gradient table");
  BV_REQUIRE(path_cells != nullptr && path_len > 0, BV_ERR_INVALID, "path_cells needs path_len > 0 slots per ray");
  return bvh::launch_train_l2_coop(s, bvh::walk_rays(s, rays), bvh::settings_of(settings), target_rgb, rgba_out, loss,
                                   bvh::GradPtrs{grads->d_density, grads->d_sh, grads->d_xyz}, 0, (hipStream_t)stream,
                                   path_cells, path_len);
}

int bv_train_step_l2_noalloc(bv_scene* s, const bv_rays* rays, const bv_trace_settings* settings,
                             const float* target_rgb, float* rgba_out, const bv_grads* grads, float* loss,
                             void* stream) {
  int rc = bvh::check_rays(s, rays);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(settings && grads, BV_ERR_INVALID, "null settings/grads");
  BV_REQUIRE(settings->eval_mode == BV_EVAL_SH, BV_ERR_UNSUPPORTED, "training differentiates SH colour only");
  BV_REQUIRE(!s->bounded, BV_ERR_UNSUPPORTED, "the adjoint supports Voronoi (radius-free) cells only");
  if (rays->count == 0) return BV_OK;
  BV_REQUIRE(rays->count < (int64_t(1) << 31), BV_ERR_INVALID, "a training call takes fewer than 2^31 rays");
  BV_REQUIRE(target_rgb != nullptr, BV_ERR_INVALID, "null target colours");
  BV_REQUIRE(grads->d_density && grads->d_sh && grads->d_xyz, BV_ERR_INVALID, "null gradient table");
  return bvh::launch_train_l2_noalloc(s, bvh::walk_rays(s, rays), bvh::settings_of(settings), target_rgb, rgba_out,
                                      loss, bvh::GradPtrs{grads->d_density, grads->d_sh, grads->d_xyz},
                                      (hipStream_t)stream);
}

int bv_l2_loss_grad(const float* rgba, const float* target_rgb, int64_t count, float* d_rgba, float* loss,
                    void* stream) {
  BV_REQUIRE(count >= 0, BV_ERR_INVALID, "negative count");
  if (count == 0) return BV_OK;
  BV_REQUIRE(rgba && target_rgb && d_rgba, BV_ERR_INVALID, "null buffer");
  hipLaunchKernelGGL(k_l2_loss_grad, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     reinterpret_cast<const float4*>(rgba), target_rgb, count, reinterpret_cast<float4*>(d_rgba),
                     loss);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

}  // extern "C"
