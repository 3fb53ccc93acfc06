// The trainer's volumetric graph over recorded paths, forward and backward in
// one kernel: blade-volume-train's build_volumetric_graph_with_options
// (diff_render.rs:943-1921) for RadFoam and PowerFoam paths without oriented
// surfaces, which meganeura evaluates as ~60 elementwise / gather / scan /
// scatter-add kernels over [P*L] tensors per training step.
//
// Per recorded step k of pixel row p (cell c_k, next cell n_k, recorded dt,
// mask):
//   sigma_k = relu(x) or softplus_beta(x), x = log_density[c_k]   (:1301-1302)
//   dt_k    = mask_k * one of
//     positions:  t_k = ((p_c + p_n) / 2 - o) . n * (n . d) / ((n . d)^2 + 1e-12),
//                 n = p_n - p_c, dt_raw = t_k - t_{k-1} (t_{-1} = 0),
//                 clamp(dt_raw, 0, 50) * gate + dt_rec * (1 - gate),
//                 gate = |n|^2 / (|n|^2 + 1e-6)                    (:1327-1473)
//     recorded:   dt_rec                                           (:1553)
//     linearized: clamp(dt_rec + sum_roles (g_role - (o, 0)) . J_role - tangent_ref, 0, 50),
//                 g = (xyz, softplus_100(log_radius)) of the previous, current
//                 and next sites                                   (:1474-1552)
//   T_k = exp(-sum_{j<k} sigma_j dt_j), alpha_k = 1 - exp(-sigma_k dt_k), w_k = T_k alpha_k
//   c_k = relu(0.5 + sum_j coef[c_k][j] basis_j)                   (:2162-2308)
//   pixel = E[view] * sum_k w_k c_k + (1 - sum_k w_k) background   (:1685-1712)
// and the loss (L1, SmoothL1, or this build's L2; plus the opacity loss).
//
// Layout: G = 16 lanes per pixel row; lane j takes step base + j of each
// 16-step chunk, so the row's gathers are one per lane and the row's
// exclusive cumsum is a group prefix scan carried across chunks.  Step k's
// dt needs t_{k-1}: a lane shuffle (the previous chunk's last t is carried).
// Pass 1 forms the pixel; pass 2 recomputes each step and applies the
// reverse mode with the suffix sums taken as (row total - inclusive prefix),
// so nothing per step is stored.  Each step's gradient values are staged in
// LDS and the group adds them step by step with lanes over the row's items
// (the SH row, density, positions, radii), so one atomic instruction covers
// one cell's contiguous SH row.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bv_host.h"

namespace {

using namespace bvk;

constexpr int kG = 16;
constexpr float kMaxPathDt = 50.0f;    // diff_render.rs:2510
constexpr float kRadiusBeta = 100.0f;  // RADIUS_SOFTPLUS_BETA, diff_render.rs:223

// diff_render.rs:2001-2022
__constant__ static const float kC0 = 0.2820948f;
__constant__ static const float kC1 = 0.4886025f;
__constant__ static const float kC2[5] = {1.0925484f, -1.0925484f, 0.31539157f, -1.0925484f, 0.5462742f};
__constant__ static const float kC3[7] = {-0.5900436f, 2.8906114f, -0.4570458f, 0.37317634f,
                                          -0.4570458f, 1.4453057f, -0.5900436f};

struct GraphArgs {
  // parameters
  const float* __restrict__ log_density;
  const float* __restrict__ positions;
  const float* __restrict__ sh;
  const float* __restrict__ log_radii;
  const float* __restrict__ exposure;
  int64_t n;
  int32_t n_views;
  // batch
  int64_t P;
  uint32_t L;
  const uint32_t* __restrict__ cells;
  const uint32_t* __restrict__ next;
  const float* __restrict__ dt_rec;
  const float* __restrict__ mask;
  const uint32_t* __restrict__ status;
  const uint32_t* __restrict__ prev;
  const float* __restrict__ ref_tangent;
  const float4* __restrict__ jac_prev;
  const float4* __restrict__ jac_cur;
  const float4* __restrict__ jac_next;
  const float* __restrict__ origin;
  const float* __restrict__ dir;
  const float* __restrict__ basis;
  const uint32_t* __restrict__ view;
  const float* __restrict__ labels;
  const float* __restrict__ target_alpha;
  // work order: the i-th row taken is row (i * row_mult) mod P, so the rows
  // the device holds at once come from all over the batch
  uint32_t row_mult;
  // options
  float beta;
  float bg[3];
  uint32_t loss;
  float opacity_weight;
  // outputs
  float* loss_out;
  float* pixels;
  float* opacity;
  float* dt_out;
  // gradients (null: not trained)
  float* d_log_density;
  float* d_positions;
  float* d_sh;
  float* d_log_radii;
  float* d_exposure;
  // when set, a cell's density and position gradients go to one (x, y, z,
  // density) slot here, added to d_positions / d_log_density after the launch
  float4* packed;
};

__device__ __forceinline__ float relu(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float step01(float x) { return x > 0.0f ? 1.0f : 0.0f; }
__device__ __forceinline__ float softplus(float x, float beta) {
  const float bx = beta * x;
  return (relu(bx) + log1pf(expf(-fabsf(bx)))) / beta;
}
__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }

// sh_basis, diff_render.rs:2062-2090 (same products, same order)
template <int K>
__device__ __forceinline__ void trainer_basis(float x, float y, float z, float (&b)[K]) {
  b[0] = kC0;
  if (K > 1) {
    b[1] = -kC1 * y;
    b[2] = kC1 * z;
    b[3] = -kC1 * x;
  }
  if (K > 4) {
    const float xx = x * x, yy = y * y, zz = z * z;
    b[4] = kC2[0] * x * y;
    b[5] = kC2[1] * y * z;
    b[6] = kC2[2] * (3.0f * zz - 1.0f);
    b[7] = kC2[3] * x * z;
    b[8] = kC2[4] * (xx - yy);
  }
  if (K > 9) {
    const float xx = x * x, yy = y * y, zz = z * z;
    b[9] = kC3[0] * y * (3.0f * xx - yy);
    b[10] = kC3[1] * x * y * z;
    b[11] = kC3[2] * y * (5.0f * zz - 1.0f);
    b[12] = kC3[3] * z * (5.0f * zz - 3.0f);
    b[13] = kC3[4] * x * (5.0f * zz - 1.0f);
    b[14] = kC3[5] * z * (xx - yy);
    b[15] = kC3[6] * x * (xx - 3.0f * yy);
  }
}

__device__ __forceinline__ float gshfl(float v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ uint32_t gshfl(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }

// Inclusive prefix sum over the group's lanes (Hillis-Steele, lanes below the
// offset keep their value).
__device__ __forceinline__ float group_scan(float v, uint32_t lane) {
#pragma unroll
  for (int off = 1; off < kG; off <<= 1) {
    const float o = __shfl_up(v, off, kG);
    v = lane >= (uint32_t)off ? v + o : v;
  }
  return v;
}
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = kG / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kG);
  return v;
}

// One step's recomputed forward values.
struct StepVals {
  uint32_t cell, next, prev;
  bool live;         // inside the row and naming cells of the model
  float mask, rec;
  // positions mode
  float t, gate, dt_raw, dt_pos, dt_clamped, nn;
  float jc[3], jn[3];  // dt/d p_cell and dt/d p_next of the crossing t_k
  float n3[3];
  // linearized mode
  float lin;
  float4 jp, jcur, jnx;
  float dt, x, dens, raw;
  float pre[3], col[3];
};

template <int K, int MODE>
__device__ __forceinline__ void step_values(const GraphArgs& a, int64_t p, uint32_t k, bool in, float3 o, float3 d,
                                            const float (&bas)[K], StepVals* s) {
  const size_t i = (size_t)p * a.L + k;
  s->cell = in ? a.cells[i] : 0u;
  s->next = in ? a.next[i] : 0u;
  s->prev = (in && MODE == BV_GRAPH_DT_LINEARIZED) ? a.prev[i] : s->cell;
  s->live = in && (int64_t)s->cell < a.n && (int64_t)s->next < a.n && (int64_t)s->prev < a.n;
  if (!s->live) s->cell = s->next = s->prev = BV_NO_CELL;  // nothing of the model is read or written
  s->mask = s->live ? a.mask[i] : 0.0f;
  s->rec = s->live ? a.dt_rec[i] : 0.0f;
  s->t = 0.0f;
  float dt_sel = s->rec;
  if (MODE == BV_GRAPH_DT_POSITIONS) {
    float pc[3] = {0.f, 0.f, 0.f}, pn[3] = {0.f, 0.f, 0.f};
    if (s->live) {
      for (int c = 0; c < 3; ++c) {
        pc[c] = a.positions[3 * (size_t)s->cell + c];
        pn[c] = a.positions[3 * (size_t)s->next + c];
      }
    }
    const float oc[3] = {o.x, o.y, o.z}, dc[3] = {d.x, d.y, d.z};
    float mo[3];
    for (int c = 0; c < 3; ++c) {
      const float mid = (pc[c] + pn[c]) * 0.5f;
      s->n3[c] = pn[c] + (-pc[c]);
      mo[c] = mid + (-oc[c]);
    }
    s->nn = (s->n3[0] * s->n3[0] + s->n3[1] * s->n3[1]) + s->n3[2] * s->n3[2];
    const float num = (mo[0] * s->n3[0] + mo[1] * s->n3[1]) + mo[2] * s->n3[2];
    const float den = (s->n3[0] * dc[0] + s->n3[1] * dc[1]) + s->n3[2] * dc[2];
    const float q = den * den + 1e-12f;
    s->t = (num * den) / q;
    // dt/dnum, dt/dden for the backward
    const float dnum = den / q;
    const float dden = num * (1e-12f - den * den) / (q * q);
    for (int c = 0; c < 3; ++c) {
      const float dn = dnum * mo[c] + dden * dc[c];
      const float dmo = dnum * s->n3[c];
      s->jc[c] = 0.5f * dmo - dn;
      s->jn[c] = 0.5f * dmo + dn;
    }
    s->gate = s->nn / (s->nn + 1e-6f);
  } else if (MODE == BV_GRAPH_DT_LINEARIZED) {
    s->jp = s->live ? a.jac_prev[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    s->jcur = s->live ? a.jac_cur[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    s->jnx = s->live ? a.jac_next[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float term[3] = {0.f, 0.f, 0.f};
    if (s->live) {
      const uint32_t roles[3] = {s->prev, s->cell, s->next};
      const float4 jac[3] = {s->jp, s->jcur, s->jnx};
      for (int r = 0; r < 3; ++r) {
        const float gx = a.positions[3 * (size_t)roles[r]] + (-o.x);
        const float gy = a.positions[3 * (size_t)roles[r] + 1] + (-o.y);
        const float gz = a.positions[3 * (size_t)roles[r] + 2] + (-o.z);
        const float gr = softplus(a.log_radii[roles[r]], kRadiusBeta) + (-0.0f);
        term[r] = ((gx * jac[r].x + gy * jac[r].y) + gz * jac[r].z) + gr * jac[r].w;
      }
    }
    const float geometry = (term[0] + term[1]) + term[2];
    const float delta = geometry + (-(s->live ? a.ref_tangent[i] : 0.0f));
    s->lin = s->rec + delta;
    const float pos = relu(s->lin);
    dt_sel = kMaxPathDt + (-relu(kMaxPathDt + (-pos)));
  }
  s->dt = dt_sel;  // completed by finish_dt (positions mode needs t_{k-1})
  s->x = s->live ? a.log_density[s->cell] : 0.0f;
  s->dens = a.beta > 0.0f ? softplus(s->x, a.beta) : relu(s->x);
  for (int c = 0; c < 3; ++c) s->pre[c] = 0.0f;
  if (s->live && s->mask != 0.0f) {
    const float* row = a.sh + (size_t)s->cell * 3 * K;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float acc = 0.0f;
#pragma unroll
      for (int j = 0; j < K; ++j) acc += row[3 * j + c] * bas[j];
      s->pre[c] = acc + 0.5f;
    }
  }
  for (int c = 0; c < 3; ++c) s->col[c] = relu(s->pre[c]);
}

template <int MODE>
__device__ __forceinline__ void finish_dt(StepVals* s, float t_prev) {
  if (MODE == BV_GRAPH_DT_POSITIONS) {
    s->dt_raw = s->t + (-t_prev);
    s->dt_pos = relu(s->dt_raw);
    s->dt_clamped = kMaxPathDt + (-relu(kMaxPathDt + (-s->dt_pos)));
    s->dt = s->dt_clamped * s->gate + s->rec * (1.0f + (-s->gate));
  }
  s->dt = s->dt * s->mask;
  s->raw = s->dens * s->dt;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Items of one step's gradient row, added by the group with lane = item:
//   [0, 3K)          SH row of the cell
//   3K               log density of the cell
//   3K+1 .. 3K+12    four (cell id, xyz) position slots
//   3K+13 .. 3K+15   log radius of the previous, current and next sites
template <int K>
struct RowItems {
  static constexpr int N = 3 * K + 16;
  static constexpr int R = (N + kG - 1) / kG;
};

template <int K, int MODE>
__global__ void __launch_bounds__(256) k_graph(GraphArgs a) {
  // per group: kG staged steps of 6 float4
  __shared__ float4 stage[256 / kG][kG][6];
  const uint32_t lane = threadIdx.x & (kG - 1);
  const uint32_t grp = threadIdx.x / kG;
  // Rows of one camera slice are neighbours in the batch and cross the same
  // cells (all start in the camera's cell), so groups run the rows in a
  // scattered order: their gradient atomics then meet on one cell far less
  // often at once.
  const int64_t i = (int64_t)blockIdx.x * (blockDim.x / kG) + grp;
  const bool valid = i < a.P;
  const int64_t p = valid ? (int64_t)(((uint64_t)i * a.row_mult) % (uint64_t)a.P) : a.P;
  const int src0 = (int)((threadIdx.x & 63u) & ~(uint32_t)(kG - 1));
  float row_loss = 0.0f;
  if (valid) {
    const float3 o = make_float3(a.origin[3 * p], a.origin[3 * p + 1], a.origin[3 * p + 2]);
    const float3 d = make_float3(a.dir[3 * p], a.dir[3 * p + 1], a.dir[3 * p + 2]);
    float bas[K];
    if (a.basis) {
      bas[0] = kC0;
#pragma unroll
      for (int j = 1; j < K; ++j) bas[j] = a.basis[(size_t)p * (K - 1) + (j - 1)];
    } else {
      trainer_basis<K>(d.x, d.y, d.z, bas);
    }
    uint32_t Lr = a.L;
    if (a.status) {
      const uint32_t cnt = a.status[p] & 0x7fffffffu;
      Lr = cnt < a.L ? cnt : a.L;
    }
    const uint32_t view = a.view ? a.view[p] : 0u;
    const bool view_ok = a.exposure != nullptr && (int64_t)view < a.n_views;
    float E[3] = {1.0f, 1.0f, 1.0f};
    if (view_ok)
      for (int c = 0; c < 3; ++c) E[c] = a.exposure[3 * (size_t)view + c];

    // ---- pass 1: forward
    float t_carry = 0.0f, cum_carry = 0.0f, P3[3] = {0.f, 0.f, 0.f}, O = 0.0f;
    for (uint32_t base = 0; base < Lr; base += kG) {
      const uint32_t k = base + lane;
      const bool in = k < Lr;
      StepVals s;
      step_values<K, MODE>(a, p, k, in, o, d, bas, &s);
      float tp = __shfl_up(s.t, 1, kG);
      tp = lane == 0 ? t_carry : tp;
      finish_dt<MODE>(&s, tp);
      const float incl = group_scan(s.raw, lane);
      float excl = __shfl_up(incl, 1, kG);
      excl = lane == 0 ? 0.0f : excl;
      const float cum = cum_carry + excl;
      const float T = expf(-cum);
      const float alpha = 1.0f + (-expf(-s.raw));
      const float w = in ? T * alpha : 0.0f;
      for (int c = 0; c < 3; ++c) P3[c] += w * s.col[c];
      O += w;
      if (a.dt_out && in) a.dt_out[(size_t)p * a.L + k] = s.dt;
      t_carry = gshfl(s.t, src0 + kG - 1);
      cum_carry += gshfl(incl, src0 + kG - 1);
    }
    if (a.dt_out)
      for (uint32_t k = Lr + lane; k < a.L; k += kG) a.dt_out[(size_t)p * a.L + k] = 0.0f;
    for (int c = 0; c < 3; ++c) P3[c] = group_sum(P3[c]);
    O = group_sum(O);

    // ---- pixel, loss and dL/dpixel
    const float inv_p = 1.0f / (float)a.P;
    float g[3];
    for (int c = 0; c < 3; ++c) {
      float pix = P3[c] * E[c];
      if (a.bg[c] != 0.0f) pix = pix + (1.0f + (-O)) * a.bg[c];
      const float diff = pix + (-a.labels[3 * p + c]);
      if (a.loss == BV_LOSS_L1) {
        row_loss += fabsf(diff) * inv_p;
        g[c] = (diff > 0.0f ? 1.0f : (diff < 0.0f ? -1.0f : 0.0f)) * inv_p;
      } else if (a.loss == BV_LOSS_SMOOTH_L1) {
        const float ab = fabsf(diff);
        const float above = relu(ab + (-1.0f));
        const float cl = ab + (-above);
        row_loss += (cl * cl * 0.5f + above) * inv_p / 3.0f;
        const float da = ab > 1.0f ? 1.0f : ab;
        g[c] = (diff > 0.0f ? da : (diff < 0.0f ? -da : 0.0f)) * inv_p / 3.0f;
      } else {
        row_loss += diff * diff * inv_p;
        g[c] = 2.0f * diff * inv_p;
      }
      if (lane == 0 && a.pixels) a.pixels[3 * p + c] = pix;
    }
    float dop = 0.0f;
    if (a.opacity_weight > 0.0f) {
      const float e = O + (-a.target_alpha[p]);
      row_loss += a.opacity_weight * e * e * inv_p;
      dop = a.opacity_weight * 2.0f * e * inv_p;
    }
    if (lane == 0 && a.opacity) a.opacity[p] = O;
    if (lane < 3 && view_ok && a.d_exposure) {
      const float v = lane == 0 ? g[0] * P3[0] : (lane == 1 ? g[1] * P3[1] : g[2] * P3[2]);
      if (v != 0.0f) atomicAdd(a.d_exposure + 3 * (size_t)view + lane, v);
    }
    // sum_k dL/dw_k w_k, with dL/dw_k = sum_c g_c (E_c c_kc - bg_c) + dop
    float gtot = dop * O;
    for (int c = 0; c < 3; ++c) gtot += g[c] * (E[c] * P3[c] - a.bg[c] * O);
    float gE[3];
    for (int c = 0; c < 3; ++c) gE[c] = g[c] * E[c];
    // the SH basis factor of each item this lane adds
    constexpr int R = RowItems<K>::R;
    float bval[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t item = lane + r * kG;
      const uint32_t j = item / 3u;
      float v = 0.0f;
#pragma unroll
      for (int q = 0; q < K; ++q) v = (j == (uint32_t)q) ? bas[q] : v;
      bval[r] = item < 3u * K ? v : 0.0f;
    }

    // ---- pass 2: reverse mode, step by step in forward order
    float ww_carry = 0.0f;
    t_carry = 0.0f;
    cum_carry = 0.0f;
    float jc_carry[3] = {0.f, 0.f, 0.f}, jn_carry[3] = {0.f, 0.f, 0.f};
    uint32_t cc_carry = BV_NO_CELL, nc_carry = BV_NO_CELL;
    float4(*st)[6] = stage[grp];
    for (uint32_t base = 0; base < Lr; base += kG) {
      const uint32_t k = base + lane;
      const bool in = k < Lr;
      StepVals s;
      step_values<K, MODE>(a, p, k, in, o, d, bas, &s);
      float tp = __shfl_up(s.t, 1, kG);
      tp = lane == 0 ? t_carry : tp;
      finish_dt<MODE>(&s, tp);
      const float incl = group_scan(s.raw, lane);
      float excl = __shfl_up(incl, 1, kG);
      excl = lane == 0 ? 0.0f : excl;
      const float cum = cum_carry + excl;
      const float T = expf(-cum);
      const float alpha = 1.0f + (-expf(-s.raw));
      const float w = in ? T * alpha : 0.0f;
      float gw = dop;
      for (int c = 0; c < 3; ++c) gw += g[c] * (E[c] * s.col[c] - a.bg[c]);
      gw = in ? gw : 0.0f;
      const float incl_ww = group_scan(gw * w, lane);
      const float suffix = gtot - (ww_carry + incl_ww);
      const float d_raw = in ? gw * (T * (1.0f - alpha)) - suffix : 0.0f;
      // density and colour
      const float d_x = d_raw * s.dt * (a.beta > 0.0f ? sigmoid(a.beta * s.x) : step01(s.x));
      float dc[3];
      for (int c = 0; c < 3; ++c) dc[c] = gE[c] * w * step01(s.pre[c]);
      const float d_sel = d_raw * s.dens * s.mask;
      // positions: slots (id, xyz) for the cell, the next cell and two more
      float g0[3] = {0.f, 0.f, 0.f}, g1[3] = {0.f, 0.f, 0.f}, g2[3] = {0.f, 0.f, 0.f}, g3[3] = {0.f, 0.f, 0.f};
      uint32_t id2 = BV_NO_CELL, id3 = BV_NO_CELL;
      float dr[3] = {0.f, 0.f, 0.f};
      // previous step's crossing Jacobian (lane - 1, or the carried last lane)
      float jcp[3], jnp[3];
      for (int c = 0; c < 3; ++c) {
        const float u = __shfl_up(s.jc[c], 1, kG), v = __shfl_up(s.jn[c], 1, kG);
        jcp[c] = lane == 0 ? jc_carry[c] : u;
        jnp[c] = lane == 0 ? jn_carry[c] : v;
      }
      uint32_t cp = __shfl_up((int)s.cell, 1, kG), np_ = __shfl_up((int)s.next, 1, kG);
      cp = lane == 0 ? cc_carry : cp;
      np_ = lane == 0 ? nc_carry : np_;
      if (MODE == BV_GRAPH_DT_POSITIONS) {
        const float d_dt_raw = d_sel * s.gate * step01(s.dt_raw) * step01(kMaxPathDt + (-s.dt_pos));
        const float d_gate = d_sel * (s.dt_clamped + (-s.rec));
        const float bb = s.nn + 1e-6f;
        const float d_nn = d_gate * 1e-6f / (bb * bb);
        for (int c = 0; c < 3; ++c) {
          const float gn = 2.0f * d_nn * s.n3[c];
          g0[c] = d_dt_raw * s.jc[c] - gn;
          g1[c] = d_dt_raw * s.jn[c] + gn;
          // t_{k-1} enters dt_raw_k with a minus sign
          g2[c] = -d_dt_raw * jcp[c];
          g3[c] = -d_dt_raw * jnp[c];
        }
        id2 = k > 0 ? cp : BV_NO_CELL;
        id3 = k > 0 ? np_ : BV_NO_CELL;
      } else if (MODE == BV_GRAPH_DT_LINEARIZED) {
        const float d_lin = d_sel * step01(s.lin) * step01(kMaxPathDt + (-relu(s.lin)));
        g0[0] = d_lin * s.jcur.x; g0[1] = d_lin * s.jcur.y; g0[2] = d_lin * s.jcur.z;
        g1[0] = d_lin * s.jnx.x; g1[1] = d_lin * s.jnx.y; g1[2] = d_lin * s.jnx.z;
        g2[0] = d_lin * s.jp.x; g2[1] = d_lin * s.jp.y; g2[2] = d_lin * s.jp.z;
        id2 = s.prev;
        if (a.d_log_radii && s.live) {
          dr[0] = d_lin * s.jp.w * sigmoid(kRadiusBeta * a.log_radii[s.prev]);
          dr[1] = d_lin * s.jcur.w * sigmoid(kRadiusBeta * a.log_radii[s.cell]);
          dr[2] = d_lin * s.jnx.w * sigmoid(kRadiusBeta * a.log_radii[s.next]);
        }
      }
      // one slot per distinct cell: fold duplicates into the first
      uint32_t id0 = s.cell, id1 = s.next;
      if (id1 == id0) { for (int c = 0; c < 3; ++c) { g0[c] += g1[c]; g1[c] = 0.f; } id1 = BV_NO_CELL; }
      if (id2 == id0) { for (int c = 0; c < 3; ++c) { g0[c] += g2[c]; g2[c] = 0.f; } id2 = BV_NO_CELL; }
      if (id2 != BV_NO_CELL && id2 == id1) { for (int c = 0; c < 3; ++c) { g1[c] += g2[c]; g2[c] = 0.f; } id2 = BV_NO_CELL; }
      if (id3 == id0) { for (int c = 0; c < 3; ++c) { g0[c] += g3[c]; g3[c] = 0.f; } id3 = BV_NO_CELL; }
      if (id3 != BV_NO_CELL && id3 == id1) { for (int c = 0; c < 3; ++c) { g1[c] += g3[c]; g3[c] = 0.f; } id3 = BV_NO_CELL; }
      if (id3 != BV_NO_CELL && id3 == id2) { for (int c = 0; c < 3; ++c) { g2[c] += g3[c]; g3[c] = 0.f; } id3 = BV_NO_CELL; }
      const bool live = in && s.live;
      st[lane][0] = make_float4(dc[0], dc[1], dc[2], d_x);
      st[lane][1] = make_float4(g0[0], g0[1], g0[2], __uint_as_float(live ? id0 : BV_NO_CELL));
      st[lane][2] = make_float4(g1[0], g1[1], g1[2], __uint_as_float(live ? id1 : BV_NO_CELL));
      st[lane][3] = make_float4(g2[0], g2[1], g2[2], __uint_as_float(live ? id2 : BV_NO_CELL));
      st[lane][4] = make_float4(g3[0], g3[1], g3[2], __uint_as_float(live ? id3 : BV_NO_CELL));
      st[lane][5] = make_float4(dr[0], dr[1], dr[2], __uint_as_float(live ? s.prev : BV_NO_CELL));
      wave_sync();
      const uint32_t cnt = Lr - base < (uint32_t)kG ? Lr - base : (uint32_t)kG;
      for (uint32_t j = 0; j < cnt; ++j) {
        const float4 A = st[j][0];
        const uint32_t cell = __float_as_uint(st[j][1].w);
        if (cell == BV_NO_CELL) continue;
        // Positions dt: step j's next cell is step j + 1's cell and step
        // j + 1's previous cell is step j's, so those two terms join the
        // cell's own before its atomics, and their own slots add nothing (3
        // position lanes a step instead of 9).  The chunk's ends add the terms
        // that reach past it.
        float4 own = st[j][1];
        bool skip_next = false, skip_prev = false;
        if (MODE == BV_GRAPH_DT_POSITIONS) {
          if (j > 0) {
            const float4 P = st[j - 1][2];
            if (__float_as_uint(P.w) == cell) { own.x += P.x; own.y += P.y; own.z += P.z; }
            skip_prev = __float_as_uint(st[j][3].w) == __float_as_uint(st[j - 1][1].w);
          }
          if (j + 1 < cnt) {
            const float4 Q = st[j + 1][3];
            if (__float_as_uint(Q.w) == cell) { own.x += Q.x; own.y += Q.y; own.z += Q.z; }
            skip_next = __float_as_uint(st[j][2].w) == __float_as_uint(st[j + 1][1].w);
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t item = lane + r * kG;
          float* addr = nullptr;
          float v = 0.0f;
          if (item < 3u * K) {
            const uint32_t c = item % 3u;
            v = (c == 0 ? A.x : (c == 1 ? A.y : A.z)) * bval[r];
            addr = a.d_sh ? a.d_sh + (size_t)cell * 3 * K + item : nullptr;
          } else if (item == 3u * K) {
            v = A.w;
            addr = a.packed ? reinterpret_cast<float*>(a.packed + cell) + 3
                            : (a.d_log_density ? a.d_log_density + cell : nullptr);
          } else if (item < 3u * K + 13u) {
            const uint32_t q = item - (3u * K + 1u);
            const uint32_t sl = q / 3u;
            // (an LDS read, then a select: selecting between `own` and the
            // LDS rows by a lane's slot index would put them in scratch)
            const float4 R = st[j][1 + sl];
            const float4 B = make_float4(sl == 0 ? own.x : R.x, sl == 0 ? own.y : R.y, sl == 0 ? own.z : R.z,
                                         sl == 0 ? own.w : R.w);
            const uint32_t c = q % 3u, id = __float_as_uint(B.w);
            const bool merged = (sl == 1 && skip_next) || (sl == 2 && skip_prev);
            v = c == 0 ? B.x : (c == 1 ? B.y : B.z);
            if (a.d_positions && id != BV_NO_CELL && !merged)
              addr = a.packed ? reinterpret_cast<float*>(a.packed + id) + c : a.d_positions + 3 * (size_t)id + c;
          } else if (item < 3u * K + 16u) {
            const uint32_t c = item - (3u * K + 13u);
            const float4 D = st[j][5];
            const uint32_t id = c == 0 ? __float_as_uint(D.w) : (c == 1 ? cell : __float_as_uint(st[j][2].w));
            // the next site's slot was folded away when it is the cell: its
            // radius term then goes with the cell's
            const uint32_t rid = (c == 2 && id == BV_NO_CELL) ? cell : id;
            v = c == 0 ? D.x : (c == 1 ? D.y : D.z);
            addr = (a.d_log_radii && rid != BV_NO_CELL) ? a.d_log_radii + rid : nullptr;
          }
          if (addr != nullptr && v != 0.0f) atomicAdd(addr, v);
        }
      }
      wave_sync();
      // carries into the next chunk
      const int last = src0 + kG - 1;
      t_carry = gshfl(s.t, last);
      cum_carry += gshfl(incl, last);
      ww_carry += gshfl(incl_ww, last);
      for (int c = 0; c < 3; ++c) {
        jc_carry[c] = gshfl(s.jc[c], last);
        jn_carry[c] = gshfl(s.jn[c], last);
      }
      cc_carry = gshfl(s.cell, last);
      nc_carry = gshfl(s.next, last);
    }
  }
  // one loss atomic per wave
  if (a.loss_out) {
    float l = lane == 0 ? row_loss : 0.0f;
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(a.loss_out, l);
  }
}

__global__ void k_graph_zero_packed(int64_t n, float4* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The packed slots added to the caller's tables (atomically: calls on other
// streams may add to the same tables).
__global__ void k_graph_unpack(int64_t n, const float4* __restrict__ p, float* d_log_density, float* d_positions) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = p[i];
  if (v.w != 0.f) atomicAdd(d_log_density + i, v.w);
  if (v.x != 0.f) atomicAdd(d_positions + 3 * i, v.x);
  if (v.y != 0.f) atomicAdd(d_positions + 3 * i + 1, v.y);
  if (v.z != 0.f) atomicAdd(d_positions + 3 * i + 2, v.z);
}

template <int K>
void launch_k(GraphArgs a, uint32_t mode, hipStream_t stream) {
  const int64_t threads = a.P * kG;
  const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
  a.row_mult = bvh::scatter_mult((uint32_t)a.P);
  switch (mode) {
    case BV_GRAPH_DT_POSITIONS: hipLaunchKernelGGL((k_graph<K, BV_GRAPH_DT_POSITIONS>), grid, block, 0, stream, a); break;
    case BV_GRAPH_DT_RECORDED: hipLaunchKernelGGL((k_graph<K, BV_GRAPH_DT_RECORDED>), grid, block, 0, stream, a); break;
    default: hipLaunchKernelGGL((k_graph<K, BV_GRAPH_DT_LINEARIZED>), grid, block, 0, stream, a); break;
  }
}

}  // namespace

extern "C" {

int bv_volume_graph_step(const bv_graph_params* params, const bv_graph_batch* batch, const bv_graph_options* options,
                         const bv_graph_outputs* outputs, const bv_graph_grads* grads, void* stream) {
  BV_REQUIRE(params && batch && options, BV_ERR_INVALID, "null params/batch/options");
  BV_REQUIRE(params->n_cells > 0 && params->n_cells < (int64_t)0xFFFFFFFF, BV_ERR_INVALID, "bad cell count");
  BV_REQUIRE(params->sh_degree >= 0 && params->sh_degree <= 3, BV_ERR_INVALID, "sh_degree must be in 0..=3");
  BV_REQUIRE(params->log_density && params->positions && params->sh, BV_ERR_INVALID, "missing parameter table");
  BV_REQUIRE(options->dt_mode <= BV_GRAPH_DT_LINEARIZED, BV_ERR_INVALID, "bad dt_mode");
  BV_REQUIRE(options->loss <= BV_LOSS_L2, BV_ERR_INVALID, "bad loss");
  BV_REQUIRE(options->softplus_beta >= 0.0f, BV_ERR_INVALID, "softplus_beta must be >= 0 (0 = ReLU)");
  BV_REQUIRE(batch->max_steps > 0, BV_ERR_INVALID, "max_steps must be positive");
  BV_REQUIRE((uint64_t)batch->num_pixels * batch->max_steps < (1ull << 40), BV_ERR_INVALID, "batch too large");
  if (batch->num_pixels == 0) return BV_OK;
  BV_REQUIRE(batch->cells && batch->next_cells && batch->dts && batch->mask, BV_ERR_INVALID, "missing path rows");
  BV_REQUIRE(batch->ray_origin && batch->ray_dir && batch->labels, BV_ERR_INVALID, "missing rays / labels");
  BV_REQUIRE(options->opacity_weight <= 0.0f || batch->target_alpha, BV_ERR_INVALID,
             "the opacity loss needs target_alpha");
  BV_REQUIRE(params->exposure == nullptr || params->n_views > 0, BV_ERR_INVALID, "exposure needs n_views > 0");
  if (options->dt_mode == BV_GRAPH_DT_LINEARIZED) {
    BV_REQUIRE(batch->previous_cells && batch->dt_reference_tangent && batch->dt_grad_previous &&
                   batch->dt_grad_current && batch->dt_grad_next,
               BV_ERR_INVALID, "the linearized dt needs the previous cells, tangents and Jacobian rows");
    BV_REQUIRE(params->log_radii, BV_ERR_INVALID, "the linearized dt needs log_radii");
  }
  GraphArgs a{};
  a.log_density = params->log_density;
  a.positions = params->positions;
  a.sh = params->sh;
  a.log_radii = params->log_radii;
  a.exposure = params->exposure;
  a.n = params->n_cells;
  a.n_views = params->n_views;
  a.P = batch->num_pixels;
  a.L = batch->max_steps;
  a.cells = batch->cells;
  a.next = batch->next_cells;
  a.dt_rec = batch->dts;
  a.mask = batch->mask;
  a.status = batch->status;
  a.prev = batch->previous_cells;
  a.ref_tangent = batch->dt_reference_tangent;
  a.jac_prev = reinterpret_cast<const float4*>(batch->dt_grad_previous);
  a.jac_cur = reinterpret_cast<const float4*>(batch->dt_grad_current);
  a.jac_next = reinterpret_cast<const float4*>(batch->dt_grad_next);
  a.origin = batch->ray_origin;
  a.dir = batch->ray_dir;
  a.basis = batch->basis;
  a.view = batch->view_idx;
  a.labels = batch->labels;
  a.target_alpha = batch->target_alpha;
  a.beta = options->softplus_beta;
  for (int c = 0; c < 3; ++c) a.bg[c] = options->background[c];
  a.loss = options->loss;
  a.opacity_weight = options->opacity_weight;
  if (outputs) {
    a.loss_out = outputs->loss;
    a.pixels = outputs->pixels;
    a.opacity = outputs->opacity;
    a.dt_out = outputs->dt;
  }
  if (grads) {
    a.d_log_density = grads->d_log_density;
    a.d_positions = options->dt_mode == BV_GRAPH_DT_RECORDED ? nullptr : grads->d_positions;
    a.d_sh = grads->d_sh;
    a.d_log_radii = options->dt_mode == BV_GRAPH_DT_LINEARIZED ? grads->d_log_radii : nullptr;
    a.d_exposure = params->exposure ? grads->d_exposure : nullptr;
  }
  hipStream_t sm = (hipStream_t)stream;
  // Density and positions through one packed (x, y, z, density) slot per
  // cell when the batch is large enough for one fewer memory line per step
  // to outweigh clearing and adding the table (the fused training kernel's
  // rule: rows at least cells / 16).
  const unsigned cgrid = (unsigned)((a.n + 255) / 256);
  if (a.d_log_density && a.d_positions && a.P * 16 >= a.n) {
    void* buf = nullptr;
    if (hipMallocAsync(&buf, sizeof(float4) * (size_t)a.n, sm) == hipSuccess) {
      a.packed = static_cast<float4*>(buf);
      hipLaunchKernelGGL(k_graph_zero_packed, dim3(cgrid), dim3(256), 0, sm, a.n, a.packed);
    } else {
      (void)hipGetLastError();  // fall back to adding straight to the tables
    }
  }
  switch (params->sh_degree) {
    case 0: launch_k<1>(a, options->dt_mode, sm); break;
    case 1: launch_k<4>(a, options->dt_mode, sm); break;
    case 2: launch_k<9>(a, options->dt_mode, sm); break;
    default: launch_k<16>(a, options->dt_mode, sm); break;
  }
  if (a.packed) {
    hipLaunchKernelGGL(k_graph_unpack, dim3(cgrid), dim3(256), 0, sm, a.n, (const float4*)a.packed,
                       a.d_log_density, a.d_positions);
    BV_HIP(hipFreeAsync(a.packed, sm));
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

}  // extern "C"
