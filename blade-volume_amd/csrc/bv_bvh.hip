// Linear BVH builder on the device (shared by the PowerFoam splat tracer and
// the Gaussian ray tracer): Morton codes of the item centres, a radix sort,
// Karras's (2012) split search for the inner nodes, a bottom-up box fit with
// release / acquire counters, and the collapse of subtrees of at most
// kLeafRange items into leaf ranges (items contiguous in Morton order).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "bv_bvh.h"
#include "bv_host.h"

namespace bvk {

__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void k_bounds_init(uint32_t* bounds) {
  if (threadIdx.x < 3) bounds[threadIdx.x] = 0xFFFFFFFFu;   // min (ordered)
  else if (threadIdx.x < 6) bounds[threadIdx.x] = 0u;        // max (ordered)
}

__global__ void k_bounds(int64_t n, const float4* __restrict__ pts, uint32_t* bounds) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float4 p = pts[i];
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float a = lo[k], b = hi[k];
    for (int m = 32; m >= 1; m >>= 1) {
      a = fminf(a, __shfl_xor(a, m));
      b = fmaxf(b, __shfl_xor(b, m));
    }
    if ((threadIdx.x & 63) == 0 && a <= b) {
      atomicMin(&bounds[k], ordered_bits(a));
      atomicMax(&bounds[3 + k], ordered_bits(b));
    }
  }
}

__global__ void k_morton(int64_t n, const float4* __restrict__ pts, const uint32_t* __restrict__ bounds,
                         uint32_t* codes, uint32_t* ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pts[i];
  const float c[3] = {p.x, p.y, p.z};
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float lo = from_ordered(bounds[k]), hi = from_ordered(bounds[3 + k]);
    const float ext = hi - lo;
    float u = ext > 0.0f ? (c[k] - lo) / ext : 0.0f;
    u = fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    q[k] = (uint32_t)u;
  }
  codes[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
  ids[i] = (uint32_t)i;
}

// Length of the common prefix of keys i and j (Morton code, then position as
// the tie-break), -1 outside [0, n).
__device__ __forceinline__ int delta(const uint32_t* __restrict__ k, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = k[i], b = k[j];
  if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(a ^ b);
}

// Karras (2012) split search: inner node i covers [min(i, j), max(i, j)].
__global__ void k_karras(int64_t n, const uint32_t* __restrict__ keys, BvhNode* nodes, uint32_t* leaf_parent,
                         uint32_t* node_lo, uint32_t* node_hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(keys, n, i, i - d);
  int64_t lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = delta(keys, n, i, j);
  int64_t s = 0;
  int64_t div = 2;
  for (;;) {
    const int64_t t = (l + div - 1) / div;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
    div <<= 1;
  }
  const int64_t g = i + s * d + (d < 0 ? -1 : 0);
  const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
  node_lo[i] = (uint32_t)lo;
  node_hi[i] = (uint32_t)hi;
  const uint32_t left = (lo == g) ? (kLeaf | (uint32_t)g) : (uint32_t)g;
  const uint32_t right = (hi == g + 1) ? (kLeaf | (uint32_t)(g + 1)) : (uint32_t)(g + 1);
  nodes[i].d.x = left;   // d.z (the parent) belongs to the parent's thread
  nodes[i].d.y = right;
  if (left & kLeaf) leaf_parent[g] = (uint32_t)i; else nodes[g].d.z = (uint32_t)i;
  if (right & kLeaf) leaf_parent[g + 1] = (uint32_t)i; else nodes[g + 1].d.z = (uint32_t)i;
}

__global__ void k_root_parent(BvhNode* nodes) { nodes[0].d.z = kNone; }

__device__ __forceinline__ void item_box(const float4* __restrict__ box_lo, const float4* __restrict__ box_hi,
                                         uint32_t item, float lo[3], float hi[3]) {
  const float4 a = box_lo[item], b = box_hi[item];
  lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
  hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
}

__device__ __forceinline__ void store_box(BvhNode* nd, bool right, const float lo[3], const float hi[3]) {
  float* a = reinterpret_cast<float*>(nd);
  const int o = right ? 6 : 0;
  __hip_atomic_store(a + o + 0, lo[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(a + o + 1, lo[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(a + o + 2, lo[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(a + o + 3, hi[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(a + o + 4, hi[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(a + o + 5, hi[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bottom-up fit: each leaf climbs; the second arrival at a node (release /
// acquire through the node's counter) merges both child boxes and goes on.
__global__ void k_fit(int64_t n, const uint32_t* __restrict__ perm, const float4* __restrict__ box_lo,
                      const float4* __restrict__ box_hi, const uint32_t* __restrict__ leaf_parent, BvhNode* nodes,
                      uint32_t* flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float lo[3], hi[3];
  item_box(box_lo, box_hi, perm[i], lo, hi);
  uint32_t child = kLeaf | (uint32_t)i;
  uint32_t p = leaf_parent[i];
  // depth <= kStack for a well-formed tree; the bounds keep a corrupt one finite
  for (int level = 0; level < kStack && p < (uint32_t)(n - 1); ++level) {
    BvhNode* nd = nodes + p;
    const bool right = __hip_atomic_load(&nd->d.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == child;
    store_box(nd, right, lo, hi);
    const uint32_t prev = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == 0u) return;
    const float* a = reinterpret_cast<const float*>(nd);
    float v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = __hip_atomic_load(a + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fminf(v[k], v[6 + k]);
      hi[k] = fmaxf(v[3 + k], v[9 + k]);
    }
    child = p;
    p = __hip_atomic_load(&nd->d.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Depth of the tree: every leaf climbs to the root and the longest path (in
// inner nodes) goes to *depth.  A near-first traversal holds at most depth + 1
// stack entries.  The climb is bounded so that a malformed tree still ends.
__global__ void k_depth(int64_t n, const uint32_t* __restrict__ leaf_parent, const BvhNode* __restrict__ nodes,
                        uint32_t* depth) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t p = leaf_parent[i], levels = 0;
  while (p < (uint32_t)(n - 1) && levels < 4096u) {
    ++levels;
    p = nodes[p].d.z;
  }
  atomicMax(depth, p < (uint32_t)(n - 1) ? 0xFFFFFFFFu : levels);
}

// Replace references to inner nodes that cover at most kLeafRange leaves by a
// leaf range (their boxes already are the union of the range's spheres).
__global__ void k_collapse(int64_t n, BvhNode* nodes, const uint32_t* __restrict__ node_lo,
                           const uint32_t* __restrict__ node_hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  uint32_t refs[2] = {nodes[i].d.x, nodes[i].d.y};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t r = refs[k];
    if (r & kLeaf) continue;
    const uint32_t size = node_hi[r] - node_lo[r] + 1u;
    if (size <= (uint32_t)kLeafRange) refs[k] = kLeaf | ((size - 1u) << 29) | node_lo[r];
  }
  nodes[i].d.x = refs[0];
  nodes[i].d.y = refs[1];
}

// One site: a root whose left child is the leaf and whose right box is empty.
__global__ void k_single(const float4* __restrict__ box_lo, const float4* __restrict__ box_hi, BvhNode* nodes) {
  float lo[3], hi[3];
  item_box(box_lo, box_hi, 0u, lo, hi);
  BvhNode nd;
  nd.a = make_float4(lo[0], lo[1], lo[2], hi[0]);
  nd.b = make_float4(hi[1], hi[2], lo[0], lo[1]);   // right slot repeats the leaf; traversal skips it
  nd.c = make_float4(lo[2], hi[0], hi[1], hi[2]);
  nd.d = make_uint4(kLeaf | 0u, kLeaf | 0u, kNone, 0u);
  nodes[0] = nd;
}

__global__ void k_fill_none(int64_t n, uint32_t* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = kNone;
}

__global__ void k_zero_u32(int64_t n, uint32_t* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace bvk

namespace bvh {

int build_lbvh(int64_t n, const float4* centres, const float4* box_lo, const float4* box_hi, bvk::BvhNode* nodes,
               uint32_t* perm, hipStream_t stream, uint32_t* depth_out) {
  BV_REQUIRE(n > 0, BV_ERR_INVALID, "BVH needs at least one item");
  BV_REQUIRE(n < (int64_t)bvk::kLeafPos, BV_ERR_INVALID, "too many items for the BVH encoding");
  const int bs = 256;
  const unsigned grid = (unsigned)((n + bs - 1) / bs);
  uint32_t *bounds = nullptr, *codes = nullptr, *ids = nullptr, *codes2 = nullptr, *leaf_parent = nullptr,
           *flags = nullptr, *node_lo = nullptr, *node_hi = nullptr, *depth = nullptr;
  uint32_t depth_h = 1;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = BV_OK;
  auto fail = [&](hipError_t e, const char* what) {
    bvh::set_error(std::string(what) + " failed: " + hipGetErrorString(e));
    rc = BV_ERR_HIP;
  };
  const size_t inner = (size_t)std::max<int64_t>(n - 1, 1);
  hipError_t e;
  do {
    if ((e = hipMalloc((void**)&bounds, 6 * 4)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&codes, 4 * (size_t)n)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&ids, 4 * (size_t)n)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&codes2, 4 * (size_t)n)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&leaf_parent, 4 * (size_t)n)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&flags, 4 * inner)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&node_lo, 4 * inner)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&node_hi, 4 * inner)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipMalloc((void**)&depth, 4)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    hipLaunchKernelGGL(bvk::k_bounds_init, dim3(1), dim3(64), 0, stream, bounds);
    const unsigned bgrid = (unsigned)std::min<int64_t>((n + bs - 1) / bs, 2048);
    hipLaunchKernelGGL(bvk::k_bounds, dim3(bgrid), dim3(bs), 0, stream, n, centres, bounds);
    hipLaunchKernelGGL(bvk::k_morton, dim3(grid), dim3(bs), 0, stream, n, centres, bounds, codes, ids);
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, codes, codes2, ids, perm, (int)n, 0, 30, stream)) !=
        hipSuccess) { fail(e, "hipcub sort (size)"); break; }
    if ((e = hipMalloc(&tmp, tmp_bytes)) != hipSuccess) { fail(e, "hipMalloc"); break; }
    if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, codes, codes2, ids, perm, (int)n, 0, 30, stream)) !=
        hipSuccess) { fail(e, "hipcub sort"); break; }
    if (n == 1) {
      hipLaunchKernelGGL(bvk::k_single, dim3(1), dim3(1), 0, stream, box_lo, box_hi, nodes);
    } else {
      const unsigned igrid = (unsigned)((n - 1 + bs - 1) / bs);
      hipLaunchKernelGGL(bvk::k_zero_u32, dim3(igrid), dim3(bs), 0, stream, n - 1, flags);
      hipLaunchKernelGGL(bvk::k_fill_none, dim3(grid), dim3(bs), 0, stream, n, leaf_parent);
      hipLaunchKernelGGL(bvk::k_karras, dim3(igrid), dim3(bs), 0, stream, n, codes2, nodes, leaf_parent, node_lo,
                         node_hi);
      hipLaunchKernelGGL(bvk::k_root_parent, dim3(1), dim3(1), 0, stream, nodes);
      hipLaunchKernelGGL(bvk::k_zero_u32, dim3(1), dim3(64), 0, stream, (int64_t)1, depth);
      hipLaunchKernelGGL(bvk::k_depth, dim3(grid), dim3(bs), 0, stream, n, leaf_parent, nodes, depth);
      hipLaunchKernelGGL(bvk::k_fit, dim3(grid), dim3(bs), 0, stream, n, perm, box_lo, box_hi, leaf_parent, nodes,
                         flags);
      hipLaunchKernelGGL(bvk::k_collapse, dim3(igrid), dim3(bs), 0, stream, n, nodes, node_lo, node_hi);
    }
    if ((e = hipGetLastError()) != hipSuccess) { fail(e, "BVH build launch"); break; }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) { fail(e, "BVH build"); break; }
    if (n > 1 && (e = hipMemcpy(&depth_h, depth, 4, hipMemcpyDeviceToHost)) != hipSuccess) {
      fail(e, "BVH depth readback");
      break;
    }
    // the traversals' LDS stacks (and the fit's climb) hold kStack entries
    if (depth_h + 1u > (uint32_t)bvk::kStack) {
      bvh::set_error("BVH depth " + std::to_string(depth_h) + " exceeds the traversal stack (" +
                     std::to_string(bvk::kStack - 1) + " levels)");
      rc = BV_ERR_UNSUPPORTED;
      break;
    }
  } while (false);
  if (depth_out) *depth_out = depth_h;
  void* temps[] = {bounds, codes, ids, codes2, leaf_parent, flags, node_lo, node_hi, depth, tmp};
  for (void* p : temps)
    if (p) (void)hipFree(p);
  return rc;
}

}  // namespace bvh
