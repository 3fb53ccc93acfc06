// Device pieces shared by the BVH traversals of the PowerFoam splat tracer and
// the Gaussian ray tracer: reference encoding, the per-lane LDS stack size, the
// slab test with pruning window, and the sorted k-buffer of (t, id) keys.
#pragma once

#include "bv_internal.h"

namespace bvk {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// Leaf reference: kLeaf | (count - 1) << 29 | first leaf position (count 1..4).
constexpr uint32_t kLeafPos = 0x1FFFFFFFu;
constexpr int kLeafRange = 4;
constexpr int kSplatK = 16;      // k-buffer depth (results do not depend on it)
constexpr int kStack = 64;       // LBVH depth <= 30 Morton bits + 32 index bits + 1
constexpr int kSplatBlock = 64;  // one wave per block: the LDS stack is per lane

// Slab test of one box against the ray; [lo_t, hi_t] is the window that can
// still hold a useful interval (already padded by the caller).
__device__ __forceinline__ bool box_hit(float x0, float y0, float z0, float x1, float y1, float z1, V3 o, V3 inv,
                                        float lo_t, float hi_t, float* tn_out) {
  const float ax = (x0 - o.x) * inv.x, bx = (x1 - o.x) * inv.x;
  const float ay = (y0 - o.y) * inv.y, by = (y1 - o.y) * inv.y;
  const float az = (z0 - o.z) * inv.z, bz = (z1 - o.z) * inv.z;
  const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), lo_t));
  const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), hi_t));
  *tn_out = tn;
  return tn <= tf;
}

// Pruning slack: sphere roots of a grazing hit may sit ~sqrt(eps) |o - c| from
// their exact value, so the box window is widened by 1e-3 (|t| + 1).
__device__ __forceinline__ float slack(float t) { return 1e-3f * (fabsf(t) + 1.0f); }

// Sorted buffer of the smallest (near, cell) keys of a traversal round.  The
// capacity is 4 in a ray's first round (most rays stop after a few intervals)
// and kSplatK after; slots past the capacity are scratch of the insertion
// network and never read.
struct KBuf {
  float t[kSplatK];
  float dt[kSplatK];
  uint32_t c[kSplatK];
  uint32_t count;
  bool small;  // capacity 4 (else kSplatK)

  __device__ __forceinline__ void clear(bool first_round) {
#pragma unroll
    for (int i = 0; i < kSplatK; ++i) {
      t[i] = INFINITY;
      dt[i] = 0.0f;
      c[i] = kNone;
    }
    count = 0;
    small = first_round;
  }
  __device__ __forceinline__ uint32_t cap() const { return small ? 4u : (uint32_t)kSplatK; }
  __device__ __forceinline__ bool full() const { return count == cap(); }
  // The largest key held when full.
  __device__ __forceinline__ float kth_t() const { return small ? t[3] : t[kSplatK - 1]; }
  __device__ __forceinline__ uint32_t kth_c() const { return small ? c[3] : c[kSplatK - 1]; }
  // Drop the first entry (static register moves; count is left to the caller).
  __device__ __forceinline__ void pop_front() {
#pragma unroll
    for (int i = 0; i + 1 < kSplatK; ++i) {
      t[i] = t[i + 1];
      dt[i] = dt[i + 1];
      c[i] = c[i + 1];
    }
    t[kSplatK - 1] = INFINITY;
    dt[kSplatK - 1] = 0.0f;
    c[kSplatK - 1] = kNone;
  }
  // Sorted insert of (tn, cell); the caller has checked that it beats the last key.
  __device__ __forceinline__ void insert(float tn, uint32_t cell, float d) {
#pragma unroll
    for (int i = kSplatK - 1; i >= 0; --i) {
      const bool before_prev = i > 0 && (tn < t[i > 0 ? i - 1 : 0] ||
                                         (tn == t[i > 0 ? i - 1 : 0] && cell < c[i > 0 ? i - 1 : 0]));
      const bool before_cur = tn < t[i] || (tn == t[i] && cell < c[i]);
      if (before_prev) {
        t[i] = t[i - (i > 0)];
        dt[i] = dt[i - (i > 0)];
        c[i] = c[i - (i > 0)];
      } else if (before_cur) {
        t[i] = tn;
        dt[i] = d;
        c[i] = cell;
      }
    }
    count = count + (count < cap() ? 1u : 0u);
  }
};

}  // namespace bvk
