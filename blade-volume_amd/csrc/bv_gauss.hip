// Gaussian ray tracing (the 3DGRT blend of trace_gaussians, trace.rs:695-773;
// GPU reference gaussian_trace.wgsl:66-177 over hardware ray queries) and its
// adjoint.
//
// Each particle responds once per ray, at its maximum-response depth
// t* = -(o'.d') / |d'|^2 in its own frame (o' = R^-1 (o - mu) / s, d' = R^-1 d / s,
// R^-1 the conjugate quaternion applied with glam's Quat * Vec3 formula), with
// alpha = opacity exp(-|o' + t* d'|^2 / 2).  Responses with opacity above
// min_opacity, t* in (t_start, t_end) and alpha >= min_opacity composite in
// (t*, index) order until transmittance <= min_transmittance; the colour is
// max(0, 0.5 + SH(d)).
//
// Design on gfx950 (no ray-tracing cores): a linear BVH over each particle's
// support ellipsoid at the scene's min_opacity (the points where alpha can
// reach min_opacity), built on the device; one lane per ray walks it with a
// flat loop (pop / node / one particle per iteration) into a k-buffer of
// (t*, index) keys with a cursor, so the blend order is exact and independent
// of K.  The particle's point of maximum response lies inside its support,
// hence inside its box: t* is bounded by the box's entry and exit depths,
// which is what prunes the traversal.  Rounds are wave-synchronous: the wave
// composites (forward) or replays (adjoint) its lanes' buffered responses
// together.  The adjoint replays the forward order with the forward colour and
// transmittance and scatters FP32 gradients of mean, scale, quaternion,
// opacity and SH.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "bv_bvh.h"
#include "bv_host.h"

struct bv_gaussians {
  int device = 0;
  int64_t n = 0;
  int32_t sh_degree = 0;
  int32_t K = 1;
  float min_opacity = 0.01f;
  int64_t device_bytes = 0;
  uint32_t depth = 0;  // longest root-to-leaf path of the BVH (inner nodes)
  // parameter masters (SoA, original order)
  float* mean = nullptr;     // [n][3]
  float* opacity = nullptr;  // [n]
  float* rot = nullptr;      // [n][4] x y z w
  float* scale = nullptr;    // [n][3]
  float* sh = nullptr;       // [n][3K]
  // BVH
  bvk::BvhNode* nodes = nullptr;  // [max(n-1,1)]
  void* leaf = nullptr;           // [n] GaussRec in Morton order
  uint32_t* perm = nullptr;       // [n]
  std::mutex mu;
};

namespace bvk {

// One particle as a leaf sees it: mean and opacity, quaternion, scale and the
// particle index (bits) in 48 bytes.
struct __align__(16) GaussRec {
  float4 a;  // mean xyz, opacity
  float4 q;  // quaternion x y z w
  float4 s;  // scale xyz, index bits
};

struct GaussView {
  const BvhNode* __restrict__ nodes;
  const GaussRec* __restrict__ leaf;
  const float* __restrict__ mean;
  const float* __restrict__ opacity;
  const float* __restrict__ rot;
  const float* __restrict__ scale;
  const float* __restrict__ sh;
  int64_t n;
  int32_t K;
  uint32_t stack_rows;  // traversal stack entries per lane: tree depth + 1
};

struct GaussSettings {
  float min_opacity, min_transmittance, t_start, t_end;
};

struct GaussGrads {
  float* mean;     // [n][3]
  float* scale;    // [n][3]
  float* rot;      // [n][4]
  float* opacity;  // [n]
  float* sh;       // [n][3K]
};

// glam's Quat * Vec3 for q = (x, y, z, w): v (w^2 - b.b) + b (2 (v.b)) + (b x v) (2 w),
// summed left to right.
__device__ __forceinline__ V3 quat_rotate(float qx, float qy, float qz, float qw, V3 v) {
  const V3 b = mk(qx, qy, qz);
  const float b2 = dot(b, b);
  const float vb = dot(v, b);
  const V3 cr = cross(b, v);
  const float s0 = qw * qw - b2, s1 = vb * 2.0f, s2 = qw * 2.0f;
  return mk((v.x * s0 + b.x * s1) + cr.x * s2, (v.y * s0 + b.y * s1) + cr.y * s2, (v.z * s0 + b.z * s1) + cr.z * s2);
}

// Support box of one particle: the ellipsoid where alpha can reach min_opacity,
// |R^-1 (x - mu) / s| <= m with m^2 = 2 ln(opacity / min_opacity), padded.
__global__ void k_gauss_prep(int64_t n, const float* __restrict__ mean, const float* __restrict__ opacity,
                             const float* __restrict__ rot, const float* __restrict__ scale, float min_opacity,
                             float4* centres, float4* box_lo, float4* box_hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float mx = mean[3 * i], my = mean[3 * i + 1], mz = mean[3 * i + 2];
  centres[i] = make_float4(mx, my, mz, 0.0f);
  const float op = opacity[i];
  const float pad = 1e-6f * fmaxf(1.0f, fmaxf(fabsf(mx), fmaxf(fabsf(my), fabsf(mz))));
  float ex = 0.f, ey = 0.f, ez = 0.f;
  if (op > min_opacity && isfinite(op)) {
    const float m = sqrtf(2.0f * logf(op / min_opacity)) * 1.001f + 1e-6f;
    const float qx = rot[4 * i], qy = rot[4 * i + 1], qz = rot[4 * i + 2], qw = rot[4 * i + 3];
    const float q2 = qx * qx + qy * qy + qz * qz + qw * qw;
    // the frame scales lengths by |q|^2, so the world semi-axes are s / |q|^2
    const float inv_q2 = q2 > 0.f ? 1.0f / q2 : 0.f;
    const float s[3] = {scale[3 * i] * inv_q2, scale[3 * i + 1] * inv_q2, scale[3 * i + 2] * inv_q2};
    float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const V3 e = mk(j == 0 ? 1.f : 0.f, j == 1 ? 1.f : 0.f, j == 2 ? 1.f : 0.f);
      const V3 c = quat_rotate(qx, qy, qz, qw, e);  // |q|^2 times column j of the rotation
      const float rx = c.x * inv_q2, ry = c.y * inv_q2, rz = c.z * inv_q2;
      acc[0] += (rx * s[j]) * (rx * s[j]);
      acc[1] += (ry * s[j]) * (ry * s[j]);
      acc[2] += (rz * s[j]) * (rz * s[j]);
    }
    ex = m * sqrtf(acc[0]) * 1.001f;
    ey = m * sqrtf(acc[1]) * 1.001f;
    ez = m * sqrtf(acc[2]) * 1.001f;
  }
  box_lo[i] = make_float4(mx - ex - pad, my - ey - pad, mz - ez - pad, 0.f);
  box_hi[i] = make_float4(mx + ex + pad, my + ey + pad, mz + ez + pad, 0.f);
}

__global__ void k_gauss_leaves(int64_t n, const uint32_t* __restrict__ perm, const float* __restrict__ mean,
                               const float* __restrict__ opacity, const float* __restrict__ rot,
                               const float* __restrict__ scale, GaussRec* leaf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = perm[i];
  GaussRec r;
  r.a = make_float4(mean[3 * k], mean[3 * k + 1], mean[3 * k + 2], opacity[k]);
  r.q = make_float4(rot[4 * k], rot[4 * k + 1], rot[4 * k + 2], rot[4 * k + 3]);
  r.s = make_float4(scale[3 * k], scale[3 * k + 1], scale[3 * k + 2], __uint_as_float(k));
  leaf[i] = r;
}

// The particle's frame of one ray: o' and d' (divided by the scales).
__device__ __forceinline__ void gauss_frame(const float4& a, const float4& q, const float4& s, V3 o, V3 d, V3* lo,
                                            V3* ld) {
  const V3 po = quat_rotate(-q.x, -q.y, -q.z, q.w, sub(o, mk(a.x, a.y, a.z)));
  const V3 pd = quat_rotate(-q.x, -q.y, -q.z, q.w, d);
  *lo = mk(po.x / s.x, po.y / s.y, po.z / s.z);
  *ld = mk(pd.x / s.x, pd.y / s.y, pd.z / s.z);
}

struct GaussCounters {
  uint32_t nodes, leaves, rounds;
};

// Traversal state of one lane: the ray, the pruning window, the cursor (keys
// already composited are <= cursor), the k-buffer of the current round, the
// LDS stack depth and the leaf range being tested.
//
// Key limit.  Compositing stops before a response once the transmittance is
// <= min_transmittance.  While the buffer is not full, `prod` is the product
// of its entries' (1 - alpha), so t_round * prod is the transmittance after
// its last entry (t_round: the transmittance at the round's start, formed as
// the compositing forms it).  Once that is <= min_transmittance (less a
// relative margin far above the rounding of the two products), compositing
// ends at or before the buffer's last key, and that key bounds the window
// instead of the K-th key of a full buffer.  An eviction from a full buffer
// drops the bound back to the K-th key.
struct GaussTrav {
  V3 o, d, inv;
  float t_hi_base, lo_base;
  uint64_t cursor;
  KBufT<false> kb;  // keys only: alpha is formed again when a response is composited
  int sp;
  uint32_t lpos, lend;
  float stop_bound, t_round, prod;
  uint64_t last;
  bool ends;
  __device__ __forceinline__ uint64_t limit() const { return ends ? last : kb.kth(); }
};

__device__ __forceinline__ void gauss_trav_init(GaussTrav& tv, V3 o, V3 d, const GaussSettings& st,
                                                uint32_t (*stk)[kSplatBlock], bool first_round_small) {
  tv.o = o;
  tv.d = d;
  tv.inv = mk(1.0f / (d.x != 0.0f ? d.x : 1e-30f), 1.0f / (d.y != 0.0f ? d.y : 1e-30f),
              1.0f / (d.z != 0.0f ? d.z : 1e-30f));
  tv.t_hi_base = st.t_end + slack(st.t_end);
  tv.lo_base = st.t_start - slack(st.t_start);
  tv.cursor = 0;
  tv.kb.clear(first_round_small);
  tv.stop_bound = st.min_transmittance * (1.0f - 1e-4f);
  tv.t_round = 1.0f;
  tv.prod = 1.0f;
  tv.last = 0;
  tv.ends = false;
  tv.sp = 0;
  stk[tv.sp++][threadIdx.x] = 0u;
  tv.lpos = 0;
  tv.lend = 0;
}

// The next round starts above the last key of a full buffer; t_round is the
// transmittance after the responses composited so far.
__device__ __forceinline__ void gauss_next_round(GaussTrav& tv, uint64_t last, float t_round,
                                                 uint32_t (*stk)[kSplatBlock]) {
  tv.cursor = last;
  const float ct = key_t(last);
  tv.lo_base = ct - slack(ct);
  tv.kb.clear(false);
  tv.t_round = t_round;
  tv.prod = 1.0f;
  tv.last = 0;
  tv.ends = false;
  stk[tv.sp++][threadIdx.x] = 0u;
}

__device__ __forceinline__ bool gauss_round_done(const GaussTrav& tv) { return tv.lpos == tv.lend && tv.sp == 0; }

// Pop one reference: a leaf range becomes the particles to test, an inner
// node's two boxes are tested and the children hit are pushed.
__device__ __forceinline__ void gauss_pop(const GaussView& g, GaussTrav& tv, uint32_t (*stk)[kSplatBlock],
                                          GaussCounters& ctr) {
  const int64_t n = g.n;
  const uint32_t lane = threadIdx.x;
  const uint32_t ref = stk[--tv.sp][lane];
  if (ref & kLeaf) {
    tv.lpos = ref & kLeafPos;
    tv.lend = tv.lpos + ((ref >> 29) & 3u) + 1u;
    if (n == 1 || (uint64_t)tv.lend > (uint64_t)n) tv.lend = tv.lpos + (tv.lpos < (uint64_t)n ? 1u : 0u);
  } else if (ref < (uint32_t)(n > 1 ? n - 1 : 1)) {
    const BvhNode nd = g.nodes[ref];
    ++ctr.nodes;
    const float kt = key_t(tv.limit());
    const float hi_t = fminf(tv.t_hi_base, kt + slack(kt));
    float tl, tr;
    const bool hl = box_hit(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, tv.o, tv.inv, tv.lo_base, hi_t, &tl);
    const bool hr =
        n != 1 && box_hit(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, tv.o, tv.inv, tv.lo_base, hi_t, &tr);
    // push the far child first so the near one is taken next
    const bool near_right = hl && hr && tr < tl;
    const uint32_t first = near_right ? nd.d.x : nd.d.y;
    const uint32_t second = near_right ? nd.d.y : nd.d.x;
    const bool first_hit = near_right ? hl : hr;
    const bool second_hit = near_right ? hr : hl;
    if (first_hit && tv.sp < (int)g.stack_rows) stk[tv.sp++][lane] = first;
    if (second_hit && tv.sp < (int)g.stack_rows) stk[tv.sp++][lane] = second;
  }
}

// Test the next particle of the leaf range.
__device__ __forceinline__ void gauss_leaf(const GaussView& g, const GaussSettings& st, GaussTrav& tv,
                                           GaussCounters& ctr) {
  const GaussRec r = g.leaf[tv.lpos];
  ++tv.lpos;
  ++ctr.leaves;
  const uint32_t idx = __float_as_uint(r.s.w);
  const float op = r.a.w;
  if (op > st.min_opacity) {
    V3 lo, ld;
    gauss_frame(r.a, r.q, r.s, tv.o, tv.d, &lo, &ld);
    const float den = dot(ld, ld);
    const float t = -dot(lo, ld) / den;
    const uint64_t key = sort_key(t, idx);
    if (t > st.t_start && t < st.t_end && key > tv.cursor && key < tv.limit()) {
      const V3 c = add(lo, scl(t, ld));
      const float alpha = op * expf(-0.5f * dot(c, c));
      if (alpha >= st.min_opacity) {
        if (tv.kb.full()) {
          tv.ends = false;  // the evicted entry's factor is unknown
        } else {
          tv.prod *= 1.0f - alpha;
          tv.last = key > tv.last ? key : tv.last;
          tv.ends = tv.t_round * tv.prod <= tv.stop_bound;
        }
        tv.kb.insert(key);
      }
    }
  }
}

// The response of particle p as the traversal formed it (the same operations
// on the same values, so the same bits): its frame (lo, ld), t* and the point
// c = lo + t* ld, and alpha = opacity exp(-|c|^2 / 2).
struct GaussResponse {
  float4 ma;  // mean, opacity
  float4 q;   // quaternion
  float4 sc;  // scale
  V3 lo, ld, c;
  float t, alpha;
};
__device__ __forceinline__ GaussResponse gauss_response(const GaussView& g, uint32_t p, V3 o, V3 d) {
  GaussResponse r;
  r.ma = make_float4(g.mean[3 * p], g.mean[3 * p + 1], g.mean[3 * p + 2], g.opacity[p]);
  r.q = make_float4(g.rot[4 * p], g.rot[4 * p + 1], g.rot[4 * p + 2], g.rot[4 * p + 3]);
  r.sc = make_float4(g.scale[3 * p], g.scale[3 * p + 1], g.scale[3 * p + 2], 0.f);
  gauss_frame(r.ma, r.q, r.sc, o, d, &r.lo, &r.ld);
  r.t = -dot(r.lo, r.ld) / dot(r.ld, r.ld);
  r.c = add(r.lo, scl(r.t, r.ld));
  r.alpha = r.ma.w * expf(-0.5f * dot(r.c, r.c));
  return r;
}

template <int DEG>
__device__ __forceinline__ V3 gauss_colour(const GaussView& g, uint32_t idx, V3 d, V3* unclamped) {
  constexpr int K = (DEG + 1) * (DEG + 1);
  const V3 c = eval_sh_row<DEG>(g.sh + (int64_t)idx * 3 * K, d);
  *unclamped = c;
  return clamp0(c);
}

// Forward: rgba [count][4], hits [count] (may be null), order rows [count][L]
// of the composited particle indices (may be null).
//
// Rounds are wave-synchronous: every lane traverses to the end of its round,
// then the wave composites the buffered responses together (SH colour, blend,
// order row).  Left to run free, each lane would composite its round alone
// while the other lanes of the wave wait, since lanes end rounds at
// different iterations.
template <int DEG>
__global__ __launch_bounds__(kSplatBlock, 3) void k_gauss_trace(GaussView g, RaySource rs, GaussSettings st,
                                                             float* rgba, uint32_t* hits, uint32_t* order,
                                                             uint32_t L, uint32_t* stats) {
  extern __shared__ uint32_t lds_stack[];  // g.stack_rows x kSplatBlock
  auto stk = reinterpret_cast<uint32_t (*)[kSplatBlock]>(lds_stack);
  const int64_t slot = (int64_t)blockIdx.x * kSplatBlock + threadIdx.x;
  V3 o = mk(0.f, 0.f, 0.f), d = mk(0.f, 0.f, 1.f);
  uint32_t start;
  int64_t idx = 0;
  // Every lane stays to the end: the rounds are wave-wide.
  const bool have = slot < ray_slots(rs) && fetch_ray(rs, slot, &o, &d, &start, &idx);
  bool active = have && normalise(&d) && g.n > 0;
  GaussCounters ctr{0u, 0u, 0u};
  float trans = 1.0f, r = 0.f, gg = 0.f, b = 0.f;
  uint32_t count = 0;
  GaussTrav tv;
  if (active) gauss_trav_init(tv, o, d, st, stk, true);
  for (;;) {
    if (active) {
      for (;;) {
        if (gauss_round_done(tv)) break;
        if (tv.lpos == tv.lend) gauss_pop(g, tv, stk, ctr);
        // Opaque to the optimiser: otherwise it threads the node path straight
        // back to the loop header, which recreates a nested node loop.
        __asm__ volatile("" : "+v"(tv.lpos));
        if (tv.lpos != tv.lend) gauss_leaf(g, st, tv, ctr);
      }
    }
    // every lane's round is complete: composite the buffered responses
    const uint32_t cnt = active ? tv.kb.count : 0u;
    const bool was_full = active && tv.kb.full();
    const uint64_t last = active ? tv.kb.kth() : 0ull;
    bool stopped = false;
    if (active) ctr.rounds += 1;
    for (uint32_t i = 0; i < cnt; ++i) {
      if (trans <= st.min_transmittance) {
        stopped = true;
        break;
      }
      const uint32_t p = key_id(tv.kb.k[0]);
      tv.kb.pop_front();
      const float alpha = gauss_response(g, p, o, d).alpha;
      V3 cu;
      const V3 col = gauss_colour<DEG>(g, p, d, &cu);
      const float w = alpha * trans;
      r += w * col.x;
      gg += w * col.y;
      b += w * col.z;
      trans *= 1.0f - alpha;
      if (order && count < L) order[idx * (int64_t)L + count] = p;
      count += 1;
    }
    active = was_full && !stopped;
    if (__ballot(active) == 0ull) break;
    if (active) gauss_next_round(tv, last, trans, stk);
  }
  if (!have) return;
  reinterpret_cast<float4*>(rgba)[idx] = make_float4(r, gg, b, 1.0f - trans);
  if (hits) hits[idx] = count;
  if (stats) {
    // algorithmic bytes: 64 per node, 48 per particle tested, 12K per composited SH row, 16 rgba
    constexpr int K = (DEG + 1) * (DEG + 1);
    stats[2 * idx] = 64u * ctr.nodes + 48u * ctr.leaves + 12u * K * count + 16u;
    stats[2 * idx + 1] = count;
  }
}

// Adjoint of the forward for dL/d(rgba), replaying the same response order with
// the forward colour C and transmittance T_final: per composited response i,
// dL/dalpha_i = g.(T_i c_i) - g.(C - C_<=i) / (1 - alpha_i) + g_a T_final / (1 - alpha_i),
// alpha = o exp(-q/2) with q = |o' + t* d'|^2 stationary in t* (dq/do' = 2 p,
// dq/dd' = 2 t* p), then through the scales and the quaternion rotation.
//
// Gradient scatter.  A float atomic is one memory request per lane, so one
// lane adding a response's 3K + 11 values costs 3K + 11 requests.  Instead the
// wave's rounds are synchronous: every lane traverses to the end of its round,
// then the wave replays the buffered responses together, one per lane per
// step.  Each lane stages its response's values as one row in LDS (the
// traversal stacks are empty at a round end and lend their space), and the
// wave adds row after row with lane t taking item t, so one atomic
// instruction covers one response's contiguous SH row (3 requests at K = 16)
// and its 11 scalars.  Consecutive rows of the same particle are summed in
// registers first.
template <int K>
struct GaussRow {
  static constexpr int kItems = 3 * K + 11;           // SH, opacity, scale xyz, mean xyz, quaternion xyzw
  static constexpr int kStride = (3 * K + 12 + 3) / 4 * 4;  // + particle index, float4-aligned
  static_assert(kItems <= kSplatBlock, "one response per atomic instruction");
  static_assert(kStride * kSplatBlock <= kStack * kSplatBlock, "rows fit the stack area");
};

template <int DEG>
__global__ __launch_bounds__(kSplatBlock) void k_gauss_adjoint(GaussView g, RaySource rs, GaussSettings st,
                                                               const float* __restrict__ rgba_fwd,
                                                               const float* __restrict__ d_rgba, GaussGrads gr) {
  constexpr int K = (DEG + 1) * (DEG + 1);
  using Row = GaussRow<K>;
  __shared__ uint32_t stk[kStack][kSplatBlock];  // the traversal stack, then the staged gradient rows
  float* rows = reinterpret_cast<float*>(&stk[0][0]);
  const uint32_t lane = threadIdx.x;
  // Every lane stays to the end: all of them take part in the scatter.
  const int64_t slot = (int64_t)blockIdx.x * kSplatBlock + lane;
  V3 o = mk(0.f, 0.f, 0.f), d = mk(0.f, 0.f, 1.f);
  uint32_t start;
  int64_t idx = 0;
  bool active = slot < ray_slots(rs) && g.n > 0;
  if (active) active = fetch_ray(rs, slot, &o, &d, &start, &idx);
  if (active) active = normalise(&d);
  float4 fw = make_float4(0.f, 0.f, 0.f, 0.f), gv = fw;
  if (active) {
    fw = reinterpret_cast<const float4*>(rgba_fwd)[idx];
    gv = reinterpret_cast<const float4*>(d_rgba)[idx];
  }
  const V3 C = mk(fw.x, fw.y, fw.z);
  const float t_final = 1.0f - fw.w;
  const V3 gc = mk(gv.x, gv.y, gv.z);
  float basis[16];
  sh_basis(K, d, basis);
  // the address of item `lane` of a particle: base + p * stride
  float* item_base = nullptr;
  uint32_t item_stride = 0;
  if (lane < 3u * K) {
    item_base = gr.sh + lane;
    item_stride = 3 * K;
  } else if (lane == 3u * K) {
    item_base = gr.opacity;
    item_stride = 1;
  } else if (lane < 3u * K + 4) {
    item_base = gr.scale + (lane - 3 * K - 1);
    item_stride = 3;
  } else if (lane < 3u * K + 7) {
    item_base = gr.mean + (lane - 3 * K - 4);
    item_stride = 3;
  } else if (lane < (uint32_t)Row::kItems) {
    item_base = gr.rot + (lane - 3 * K - 7);
    item_stride = 4;
  }
  GaussCounters ctr{0u, 0u, 0u};
  GaussTrav tv;
  if (active) gauss_trav_init(tv, o, d, st, stk, true);
  float Ti = 1.0f;
  V3 acc = mk(0.f, 0.f, 0.f);
  bool terminated = false;
  for (;;) {
    if (active) {
      for (;;) {
        if (gauss_round_done(tv)) break;
        if (tv.lpos == tv.lend) gauss_pop(g, tv, stk, ctr);
        __asm__ volatile("" : "+v"(tv.lpos));
        if (tv.lpos != tv.lend) gauss_leaf(g, st, tv, ctr);
      }
    }
    // every lane's round is complete: replay the buffered responses together
    const uint32_t cnt = active ? tv.kb.count : 0u;
    const bool was_full = active && tv.kb.full();
    const uint64_t last = active ? tv.kb.kth() : 0ull;
    for (uint32_t i = 0; __ballot(i < cnt) != 0ull; ++i) {
      if (i < cnt && !terminated && Ti <= st.min_transmittance) terminated = true;
      const bool has = i < cnt && !terminated;
      if (has) {
        const uint32_t p = key_id(tv.kb.k[0]);
        // the response again: frame, t*, the point c and alpha
        const GaussResponse rsp = gauss_response(g, p, o, d);
        const float a = rsp.alpha;
        float* row = rows + lane * Row::kStride;
        V3 cu;
        const V3 c = gauss_colour<DEG>(g, p, d, &cu);
        const float wgt = a * Ti;
        acc = add(acc, scl(wgt, c));
        const V3 rest = sub(C, acc);
        const float dA = dot(gc, scl(Ti, c)) - dot(gc, rest) / (1.0f - a) + gv.w * t_final / (1.0f - a);
        // SH (through the per-channel clamp)
        const float gcol[3] = {cu.x > 0.f ? gc.x * wgt : 0.f, cu.y > 0.f ? gc.y * wgt : 0.f,
                               cu.z > 0.f ? gc.z * wgt : 0.f};
#pragma unroll
        for (int k = 0; k < K; ++k) {
          row[3 * k] = gcol[0] * basis[k];
          row[3 * k + 1] = gcol[1] * basis[k];
          row[3 * k + 2] = gcol[2] * basis[k];
        }
        const float op = rsp.ma.w;
        row[3 * K] = dA * (a / op);
        const float dq = dA * a * -0.5f;
        const float4 q = rsp.q, sc = rsp.sc;
        const V3 u = sub(o, mk(rsp.ma.x, rsp.ma.y, rsp.ma.z));
        const V3 lo = rsp.lo, ld = rsp.ld, pc = rsp.c;
        const float tt = rsp.t;
        const float glo[3] = {2.0f * dq * pc.x, 2.0f * dq * pc.y, 2.0f * dq * pc.z};
        const float k2 = 2.0f * dq * tt;
        const float gld[3] = {k2 * pc.x, k2 * pc.y, k2 * pc.z};
        const float lov[3] = {lo.x, lo.y, lo.z}, ldv[3] = {ld.x, ld.y, ld.z}, sv[3] = {sc.x, sc.y, sc.z};
        float gpo[3], gpd[3];
#pragma unroll
        for (int a3 = 0; a3 < 3; ++a3) {
          row[3 * K + 1 + a3] = -(glo[a3] * lov[a3] + gld[a3] * ldv[a3]) / sv[a3];
          gpo[a3] = glo[a3] / sv[a3];
          gpd[a3] = gld[a3] / sv[a3];
        }
        // f(b, w, v) = v (w^2 - b.b) + b (2 v.b) + (b x v) (2 w) with b = -q.xyz
        const float bb[3] = {-q.x, -q.y, -q.z};
        const float w = q.w;
        const float bdotb = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
        const float s0 = w * w - bdotb;
        float gb[3] = {0.f, 0.f, 0.f}, gw = 0.f, gu[3] = {0.f, 0.f, 0.f};
        const float uu[3] = {u.x, u.y, u.z}, dd[3] = {d.x, d.y, d.z};
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          const float* vv = pass == 0 ? uu : dd;
          const float* gf = pass == 0 ? gpo : gpd;
          const float vb = vv[0] * bb[0] + vv[1] * bb[1] + vv[2] * bb[2];
          const float bxv[3] = {bb[1] * vv[2] - bb[2] * vv[1], bb[2] * vv[0] - bb[0] * vv[2],
                                bb[0] * vv[1] - bb[1] * vv[0]};
#pragma unroll
          for (int a3 = 0; a3 < 3; ++a3) {
            gw += gf[a3] * (2.0f * w * vv[a3] + 2.0f * bxv[a3]);
            const int a1 = (a3 + 1) % 3, a2 = (a3 + 2) % 3;
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3) {
              float eps_v = 0.f, eps_b = 0.f;
              if (c3 == a2) {
                eps_v = bb[a1];
                eps_b = -vv[a1];
              }
              if (c3 == a1) {
                eps_v = -bb[a2];
                eps_b = vv[a2];
              }
              const float dv = (a3 == c3 ? s0 : 0.f) + 2.0f * bb[a3] * bb[c3] + 2.0f * w * eps_v;
              const float db = -2.0f * bb[c3] * vv[a3] + (a3 == c3 ? 2.0f * vb : 0.f) + 2.0f * bb[a3] * vv[c3] +
                               2.0f * w * eps_b;
              if (pass == 0) gu[c3] += gf[a3] * dv;
              gb[c3] += gf[a3] * db;
            }
          }
        }
#pragma unroll
        for (int a3 = 0; a3 < 3; ++a3) {
          row[3 * K + 4 + a3] = -gu[a3];
          row[3 * K + 7 + a3] = -gb[a3];
        }
        row[3 * K + 10] = gw;
        row[Row::kItems] = __uint_as_float(p);
        Ti *= 1.0f - a;
      }
      if (i < cnt) tv.kb.pop_front();
      __syncthreads();
      // add the staged rows, one per atomic instruction, lane = item
      uint64_t pending = __ballot(has);
      uint32_t pcur = kNone;
      float sum = 0.0f;
      while (pending != 0ull) {
        const uint32_t src = (uint32_t)__builtin_ctzll(pending);
        pending &= pending - 1ull;
        const float* row = rows + src * Row::kStride;
        const uint32_t p = __builtin_amdgcn_readfirstlane(__float_as_uint(row[Row::kItems]));
        const float v = lane < (uint32_t)Row::kItems ? row[lane] : 0.0f;
        if (p != pcur) {
          if (pcur != kNone && item_base) atomicAdd(item_base + (size_t)pcur * item_stride, sum);
          pcur = p;
          sum = v;
        } else {
          sum += v;
        }
      }
      if (pcur != kNone && item_base) atomicAdd(item_base + (size_t)pcur * item_stride, sum);
      __syncthreads();
    }
    active = was_full && !terminated;
    if (__ballot(active) == 0ull) break;
    if (active) gauss_next_round(tv, last, Ti, stk);
  }
}

__global__ void k_gauss_zero(int64_t n, float* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

}  // namespace bvk

// ---------------------------------------------------------------- host side

namespace {

void free_gaussians(bv_gaussians* g) {
  void* ptrs[] = {g->mean, g->opacity, g->rot, g->scale, g->sh, g->nodes, g->leaf, g->perm};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete g;
}

int rebuild(bv_gaussians* g, hipStream_t stream) {
  const int64_t n = g->n;
  if (n == 0) return BV_OK;
  float4 *centres = nullptr, *lo = nullptr, *hi = nullptr;
  BV_HIP(hipMalloc((void**)&centres, 16 * (size_t)n));
  if (hipMalloc((void**)&lo, 16 * (size_t)n) != hipSuccess || hipMalloc((void**)&hi, 16 * (size_t)n) != hipSuccess) {
    (void)hipFree(centres);
    if (lo) (void)hipFree(lo);
    bvh::set_error("hipMalloc of Gaussian boxes failed");
    return BV_ERR_OOM;
  }
  const int bs = 256;
  const unsigned grid = (unsigned)((n + bs - 1) / bs);
  hipLaunchKernelGGL(bvk::k_gauss_prep, dim3(grid), dim3(bs), 0, stream, n, g->mean, g->opacity, g->rot, g->scale,
                     g->min_opacity, centres, lo, hi);
  int rc = bvh::build_lbvh(n, centres, lo, hi, g->nodes, g->perm, stream, &g->depth);
  if (rc == BV_OK) {
    hipLaunchKernelGGL(bvk::k_gauss_leaves, dim3(grid), dim3(bs), 0, stream, n, g->perm, g->mean, g->opacity, g->rot,
                       g->scale, reinterpret_cast<bvk::GaussRec*>(g->leaf));
    if (hipStreamSynchronize(stream) != hipSuccess) {
      bvh::set_error("Gaussian leaf gather failed");
      rc = BV_ERR_HIP;
    }
  }
  (void)hipFree(centres);
  (void)hipFree(lo);
  (void)hipFree(hi);
  return rc;
}

bvk::GaussView view_of(const bv_gaussians* g) {
  return bvk::GaussView{g->nodes, reinterpret_cast<const bvk::GaussRec*>(g->leaf), g->mean, g->opacity, g->rot,
                        g->scale, g->sh, g->n, g->K, bvk::stack_rows_for(g->depth)};
}

int check_settings(const bv_gaussians* g, const bv_gaussian_settings* s) {
  BV_REQUIRE(g && s, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(std::isfinite(s->min_opacity) && s->min_opacity > 0.0f, BV_ERR_INVALID,
             "min_opacity must be finite and positive");
  BV_REQUIRE(s->min_opacity >= g->min_opacity, BV_ERR_INVALID,
             "min_opacity below the support the particles were built with");
  BV_REQUIRE(std::isfinite(s->min_transmittance) && s->min_transmittance >= 0.0f && s->min_transmittance <= 1.0f,
             BV_ERR_INVALID, "min_transmittance must be in [0, 1]");
  BV_REQUIRE(std::isfinite(s->t_start) && std::isfinite(s->t_end) && s->t_end > s->t_start, BV_ERR_INVALID,
             "trace interval must be finite and non-empty");
  return BV_OK;
}

template <int DEG>
void launch_trace_t(const bv_gaussians* g, const bvk::RaySource& rs, const bvk::GaussSettings& st, float* rgba,
                    uint32_t* hits, uint32_t* order, uint32_t L, uint32_t* stats, hipStream_t stream) {
  const int64_t slots = bvk::ray_slots(rs);
  const unsigned grid = (unsigned)((slots + bvk::kSplatBlock - 1) / bvk::kSplatBlock);
  const bvk::GaussView v = view_of(g);
  hipLaunchKernelGGL((bvk::k_gauss_trace<DEG>), dim3(grid), dim3(bvk::kSplatBlock), bvk::stack_bytes(v.stack_rows),
                     stream, v, rs, st, rgba, hits, order, L, stats);
}

template <int DEG>
void launch_adjoint_t(const bv_gaussians* g, const bvk::RaySource& rs, const bvk::GaussSettings& st,
                      const float* rgba_fwd, const float* d_rgba, const bvk::GaussGrads& gr, hipStream_t stream) {
  const int64_t slots = bvk::ray_slots(rs);
  const unsigned grid = (unsigned)((slots + bvk::kSplatBlock - 1) / bvk::kSplatBlock);
  hipLaunchKernelGGL((bvk::k_gauss_adjoint<DEG>), dim3(grid), dim3(bvk::kSplatBlock), 0, stream, view_of(g), rs, st,
                     rgba_fwd, d_rgba, gr);
}

int trace(bv_gaussians* g, const bvk::RaySource& rs, const bv_gaussian_settings* settings,
          const bv_gaussian_outputs* out, hipStream_t stream) {
  int rc = check_settings(g, settings);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(out != nullptr && out->rgba != nullptr, BV_ERR_INVALID, "rgba output is required");
  BV_REQUIRE(out->order == nullptr || out->order_len > 0, BV_ERR_INVALID, "order rows need order_len > 0");
  if (bvk::ray_slots(rs) == 0 || g->n == 0) {
    if (g->n == 0 && bvk::ray_slots(rs) > 0) {
      BV_HIP(hipMemsetAsync(out->rgba, 0, 16 * (size_t)rs.count, stream));
      if (out->hits) BV_HIP(hipMemsetAsync(out->hits, 0, 4 * (size_t)rs.count, stream));
    }
    return BV_OK;
  }
  const bvk::GaussSettings st{settings->min_opacity, settings->min_transmittance, settings->t_start,
                              settings->t_end};
  switch (g->sh_degree < 3 ? g->sh_degree : 3) {
    case 0: launch_trace_t<0>(g, rs, st, out->rgba, out->hits, out->order, out->order_len, out->stats, stream); break;
    case 1: launch_trace_t<1>(g, rs, st, out->rgba, out->hits, out->order, out->order_len, out->stats, stream); break;
    case 2: launch_trace_t<2>(g, rs, st, out->rgba, out->hits, out->order, out->order_len, out->stats, stream); break;
    default: launch_trace_t<3>(g, rs, st, out->rgba, out->hits, out->order, out->order_len, out->stats, stream); break;
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int upload(float** dst, const float* src, size_t count) {
  BV_HIP(hipMalloc((void**)dst, 4 * (count > 0 ? count : 1)));
  if (count) BV_HIP(hipMemcpy(*dst, src, 4 * count, hipMemcpyHostToDevice));
  return BV_OK;
}

}  // namespace

extern "C" {

int bv_gaussians_create(const bv_gaussians_desc* desc, bv_gaussians** out) {
  BV_REQUIRE(desc && out, BV_ERR_INVALID, "null argument");
  *out = nullptr;
  BV_REQUIRE(desc->count >= 0 && desc->count < (int64_t)bvk::kLeafPos, BV_ERR_INVALID, "particle count out of range");
  BV_REQUIRE(desc->sh_degree >= 0 && desc->sh_degree <= 3, BV_ERR_INVALID, "sh_degree must be 0..3");
  BV_REQUIRE(std::isfinite(desc->min_opacity) && desc->min_opacity > 0.0f, BV_ERR_INVALID,
             "min_opacity must be finite and positive");
  const int64_t n = desc->count;
  BV_REQUIRE(n == 0 || (desc->means && desc->opacity && desc->rotations && desc->scales && desc->sh), BV_ERR_INVALID,
             "null particle arrays");
  for (int64_t i = 0; i < n; ++i) {
    const float* s = desc->scales + 3 * i;
    BV_REQUIRE(s[0] > 0.0f && s[1] > 0.0f && s[2] > 0.0f && std::isfinite(s[0]) && std::isfinite(s[1]) &&
                   std::isfinite(s[2]),
               BV_ERR_INVALID, "Gaussian scales must be finite and positive");
  }
  bv_gaussians* g = new bv_gaussians();
  BV_HIP(hipGetDevice(&g->device));
  g->n = n;
  g->sh_degree = desc->sh_degree;
  g->K = (desc->sh_degree + 1) * (desc->sh_degree + 1);
  g->min_opacity = desc->min_opacity;
  int rc = BV_OK;
  if ((rc = upload(&g->mean, desc->means, 3 * (size_t)n)) != BV_OK ||
      (rc = upload(&g->opacity, desc->opacity, (size_t)n)) != BV_OK ||
      (rc = upload(&g->rot, desc->rotations, 4 * (size_t)n)) != BV_OK ||
      (rc = upload(&g->scale, desc->scales, 3 * (size_t)n)) != BV_OK ||
      (rc = upload(&g->sh, desc->sh, 3 * (size_t)g->K * n)) != BV_OK) {
    free_gaussians(g);
    return rc;
  }
  if (n > 0) {
    const size_t node_bytes = (size_t)std::max<int64_t>(n - 1, 1) * sizeof(bvk::BvhNode);
    if (hipMalloc((void**)&g->nodes, node_bytes) != hipSuccess ||
        hipMalloc(&g->leaf, (size_t)n * sizeof(bvk::GaussRec)) != hipSuccess ||
        hipMalloc((void**)&g->perm, 4 * (size_t)n) != hipSuccess) {
      free_gaussians(g);
      bvh::set_error("hipMalloc of the Gaussian BVH failed");
      return BV_ERR_OOM;
    }
    g->device_bytes = (int64_t)(node_bytes + (size_t)n * (sizeof(bvk::GaussRec) + 4 + 4 * (3 + 1 + 4 + 3 + 3 * g->K)));
    rc = rebuild(g, nullptr);
    if (rc != BV_OK) {
      free_gaussians(g);
      return rc;
    }
  }
  *out = g;
  return BV_OK;
}

int bv_gaussians_destroy(bv_gaussians* g) {
  if (!g) return BV_OK;
  BV_HIP(hipDeviceSynchronize());
  free_gaussians(g);
  return BV_OK;
}

int bv_gaussians_update(bv_gaussians* g, const float* means, const float* opacity, const float* rotations,
                        const float* scales, const float* sh, void* stream) {
  BV_REQUIRE(g != nullptr, BV_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lock(g->mu);
  hipStream_t st = (hipStream_t)stream;
  const size_t n = (size_t)g->n;
  if (means) BV_HIP(hipMemcpyAsync(g->mean, means, 12 * n, hipMemcpyDeviceToDevice, st));
  if (opacity) BV_HIP(hipMemcpyAsync(g->opacity, opacity, 4 * n, hipMemcpyDeviceToDevice, st));
  if (rotations) BV_HIP(hipMemcpyAsync(g->rot, rotations, 16 * n, hipMemcpyDeviceToDevice, st));
  if (scales) BV_HIP(hipMemcpyAsync(g->scale, scales, 12 * n, hipMemcpyDeviceToDevice, st));
  if (sh) BV_HIP(hipMemcpyAsync(g->sh, sh, 12 * (size_t)g->K * n, hipMemcpyDeviceToDevice, st));
  if (means || opacity || rotations || scales) return rebuild(g, st);
  return BV_OK;
}

int bv_gaussians_trace_rays(bv_gaussians* g, const bv_rays* rays, const bv_gaussian_settings* settings,
                            const bv_gaussian_outputs* out, void* stream) {
  BV_REQUIRE(rays != nullptr, BV_ERR_INVALID, "rays is null");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(rays->count == 0 || (rays->origins && rays->directions), BV_ERR_INVALID, "null ray arrays");
  return trace(g, bvh::ray_source(rays), settings, out, (hipStream_t)stream);
}

int bv_gaussians_trace_camera(bv_gaussians* g, const bv_camera* camera, uint32_t width, uint32_t height,
                              const bv_gaussian_settings* settings, const bv_gaussian_outputs* out, void* stream) {
  BV_REQUIRE(camera != nullptr, BV_ERR_INVALID, "camera is null");
  BV_REQUIRE(width > 0 && height > 0, BV_ERR_INVALID, "resolution must be non-zero");
  bvk::RaySource rs = bvh::camera_source(camera, width, height, 0u);
  rs.count = (int64_t)width * height;
  return trace(g, rs, settings, out, (hipStream_t)stream);
}

int bv_gaussians_adjoint(bv_gaussians* g, const bv_rays* rays, const bv_gaussian_settings* settings,
                         const float* rgba_fwd, const float* d_rgba, const bv_gaussian_grads* grads, void* stream) {
  int rc = check_settings(g, settings);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(rays != nullptr && grads != nullptr, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  if (rays->count == 0 || g->n == 0) return BV_OK;
  BV_REQUIRE(rays->origins && rays->directions && rgba_fwd && d_rgba, BV_ERR_INVALID, "null ray or colour arrays");
  BV_REQUIRE(grads->d_mean && grads->d_scale && grads->d_rotation && grads->d_opacity && grads->d_sh, BV_ERR_INVALID,
             "null gradient table");
  const bvk::RaySource rs = bvh::ray_source(rays);
  const bvk::GaussSettings st{settings->min_opacity, settings->min_transmittance, settings->t_start,
                              settings->t_end};
  const bvk::GaussGrads gr{grads->d_mean, grads->d_scale, grads->d_rotation, grads->d_opacity, grads->d_sh};
  hipStream_t s = (hipStream_t)stream;
  switch (g->sh_degree < 3 ? g->sh_degree : 3) {
    case 0: launch_adjoint_t<0>(g, rs, st, rgba_fwd, d_rgba, gr, s); break;
    case 1: launch_adjoint_t<1>(g, rs, st, rgba_fwd, d_rgba, gr, s); break;
    case 2: launch_adjoint_t<2>(g, rs, st, rgba_fwd, d_rgba, gr, s); break;
    default: launch_adjoint_t<3>(g, rs, st, rgba_fwd, d_rgba, gr, s); break;
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int bv_gaussians_get_info(const bv_gaussians* g, int64_t* count, int64_t* device_bytes) {
  BV_REQUIRE(g != nullptr, BV_ERR_INVALID, "null argument");
  if (count) *count = g->n;
  if (device_bytes) *device_bytes = g->device_bytes;
  return BV_OK;
}

}  // extern "C"
