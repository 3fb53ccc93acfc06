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

// The traversal stack is dynamic LDS of stack_rows x kSplatBlock entries
// (row-major, lane-minor); stack_rows = tree depth + 1 bounds a near-first
// traversal, and a Karras tree over 30 Morton bits and at most 32 index bits
// is at most 62 deep, so kStack rows always suffice.
__host__ __device__ inline uint32_t stack_rows_for(uint32_t depth) {
  const uint32_t r = depth + 1u;
  return r < 2u ? 2u : (r > (uint32_t)kStack ? (uint32_t)kStack : r);
}
__host__ inline size_t stack_bytes(uint32_t rows) { return (size_t)rows * kSplatBlock * sizeof(uint32_t); }
// LDS of a relit splat wave's per-light colour sums: [3 BV_MAX_LIGHTS][kSplatBlock] floats.
constexpr size_t kLightSumBytes = (size_t)3 * BV_MAX_LIGHTS * kSplatBlock * sizeof(float);

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

// Box of the sphere (p.xyz, max(p.w, 0)) padded by 16 ulp of the site scale
// max(|x|, |y|, |z|, r, 1): wide enough that every sphere a rounded FP32 test
// finds touching or crossing it overlaps the box.
__device__ __forceinline__ void sphere_box(float4 p, float4* lo, float4* hi) {
  const float r = p.w > 0.0f ? p.w : 0.0f;
  const float scale = fmaxf(fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fmaxf(fabsf(p.z), r)), 1.0f);
  const float e = r + 16.0f * 1.1920929e-7f * scale;
  *lo = make_float4(p.x - e, p.y - e, p.z - e, 0.0f);
  *hi = make_float4(p.x + e, p.y + e, p.z + e, 0.0f);
}

// Pruning slack: sphere roots of a grazing hit may sit ~sqrt(eps) |o - c| from
// their exact value, so the box window is widened by 1e-3 (|t| + 1).
__device__ __forceinline__ float slack(float t) { return 1e-3f * (fabsf(t) + 1.0f); }

// (t, id) as one unsigned 64-bit key whose integer order is the lexicographic
// order of (t as a float, id): the float's bits are mapped so that unsigned
// comparison follows float comparison (-0 is folded into +0 first, as the
// float comparison treats them equal), the id fills the low word.
__device__ __forceinline__ uint64_t sort_key(float t, uint32_t id) {
  uint32_t u = __float_as_uint(t + 0.0f);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((uint64_t)u << 32) | id;
}
__device__ __forceinline__ float key_t(uint64_t k) {
  uint32_t u = (uint32_t)(k >> 32);
  u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
  return __uint_as_float(u);
}
__device__ __forceinline__ uint32_t key_id(uint64_t k) { return (uint32_t)k; }
// Above every key a traversal inserts: (+inf, 0xFFFFFFFF).
constexpr uint64_t kEmptyKey = (uint64_t(0xFF800000u) << 32) | 0xFFFFFFFFull;

// Sorted buffer of the smallest (t, id) keys of a traversal round, with one
// float payload each when PAYLOAD.  The capacity is 4 in a first round that
// asks for it and kSplatK otherwise; slots past the capacity are scratch of
// the insertion network and never read.  Empty slots hold kEmptyKey, so "not
// full, or below the last held key" is one comparison against kth().  Every
// access uses a static slot index (registers only).
template <bool PAYLOAD>
struct KBufT {
  uint64_t k[kSplatK];
  float v[PAYLOAD ? kSplatK : 1];
  uint32_t count;
  bool small;  // capacity 4 (else kSplatK)

  __device__ __forceinline__ void clear(bool first_round) {
#pragma unroll
    for (int i = 0; i < kSplatK; ++i) {
      k[i] = kEmptyKey;
      if (PAYLOAD) v[i] = 0.0f;
    }
    count = 0;
    small = first_round;
  }
  __device__ __forceinline__ uint32_t cap() const { return small ? 4u : (uint32_t)kSplatK; }
  __device__ __forceinline__ bool full() const { return count == cap(); }
  // The largest key held when full, kEmptyKey otherwise.
  __device__ __forceinline__ uint64_t kth() const { return small ? k[3] : k[kSplatK - 1]; }
  // Drop the first entry (count is left to the caller).
  __device__ __forceinline__ void pop_front() {
#pragma unroll
    for (int i = 0; i + 1 < kSplatK; ++i) {
      k[i] = k[i + 1];
      if (PAYLOAD) v[i] = v[i + 1];
    }
    k[kSplatK - 1] = kEmptyKey;
    if (PAYLOAD) v[kSplatK - 1] = 0.0f;
  }
  // Sorted insert; the caller has checked key < kth().  Slot i takes its
  // left neighbour if the key sorts before that neighbour, else the key if it
  // sorts before slot i, else keeps its entry (selects, no branches).
  __device__ __forceinline__ void insert(uint64_t key, float val = 0.0f) {
    bool before[kSplatK];
#pragma unroll
    for (int i = 0; i < kSplatK; ++i) before[i] = key < k[i];
#pragma unroll
    for (int i = kSplatK - 1; i > 0; --i) {
      const uint64_t ki = before[i] ? key : k[i];
      k[i] = before[i - 1] ? k[i - 1] : ki;
      if (PAYLOAD) {
        const float vi = before[i] ? val : v[i];
        v[i] = before[i - 1] ? v[i - 1] : vi;
      }
    }
    k[0] = before[0] ? key : k[0];
    if (PAYLOAD) v[0] = before[0] ? val : v[0];
    count = count + (count < cap() ? 1u : 0u);
  }
};
using KBuf = KBufT<true>;

}  // namespace bvk
