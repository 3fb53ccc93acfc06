// Čech adjacency of weighted sites on the device.
//
// Reference semantics: adjacency::compute_cech (blade-volume/src/adjacency.rs:
// 406-425, compute_cech_with_workers :427-540) with the default
// AdjacencyConfig, whose rows build_symmetric_csr sorts and de-duplicates
// (:55-85).  An edge {i, j}, i != j, exists when the FP32 squared distance
// ((dx dx + dy dy) + dz dz) is <= (r_i + r_j)^2, radii clamped below at 0.
// The reference finds candidates with one k-d tree per radius band; which
// structure proposes candidates does not change the rows, only the exact
// predicate does.
//
// Design on gfx950:
//   * A linear BVH over the sites' padded sphere boxes (the splat tracer's
//     builder), so a query descends only into subtrees whose spheres can
//     reach the query sphere.  The padding covers every pair that the rounded
//     predicate accepts, so the predicate alone decides each edge.
//   * One lane per site, lanes in Morton order: a wave queries 64 neighbouring
//     sites, which walk the same nodes.  The walk is one flat loop (a node pop
//     or one leaf site per iteration) with a per-lane LDS stack, as in the
//     splat traversal.
//   * Two passes of the same query: the first counts each row (and the total
//     edge count), an exclusive scan turns the counts into CSR offsets, the
//     second writes each row in leaf order; a segmented radix sort then puts
//     every row in ascending site order.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include <algorithm>
#include <cmath>
#include <string>

#include "bv_bvh.h"
#include "bv_host.h"


namespace bvk {

// (x, y, z, max(r, 0)) per site; a NaN radius counts as 0, as Rust's f32::max
// takes the other operand.
__global__ void k_cech_sites(int64_t n, const float* __restrict__ xyz, const float* __restrict__ radius,
                             float4* pts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = radius[i];
  pts[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], r > 0.0f ? r : 0.0f);
}

__global__ void k_cech_boxes(int64_t n, const float4* __restrict__ pts, float4* lo, float4* hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sphere_box(pts[i], &lo[i], &hi[i]);
}

// Sites in leaf (Morton) order.
__global__ void k_cech_leaves(int64_t n, const uint32_t* __restrict__ perm, const float4* __restrict__ pts,
                              float4* leaf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  leaf[i] = pts[perm[i]];
}

__device__ __forceinline__ bool boxes_overlap(float x0, float y0, float z0, float x1, float y1, float z1, float4 lo,
                                              float4 hi) {
  return x0 <= hi.x && y0 <= hi.y && z0 <= hi.z && x1 >= lo.x && y1 >= lo.y && z1 >= lo.z;
}

// FILL = false: counts[i] = row length of site i, *total += every row length.
// FILL = true: row i written at neighbors[offsets[i]...] in leaf order.
template <bool FILL>
__global__ __launch_bounds__(kSplatBlock) void k_cech_query(int64_t n, const BvhNode* __restrict__ nodes,
                                                            const uint32_t* __restrict__ perm,
                                                            const float4* __restrict__ leaf, uint32_t stack_rows,
                                                            uint32_t* __restrict__ counts,
                                                            unsigned long long* __restrict__ total,
                                                            const uint32_t* __restrict__ offsets,
                                                            uint32_t* __restrict__ neighbors) {
  extern __shared__ uint32_t lds_stack[];
  auto stk = reinterpret_cast<uint32_t (*)[kSplatBlock]>(lds_stack);
  const int64_t t = (int64_t)blockIdx.x * kSplatBlock + threadIdx.x;
  uint32_t cnt = 0;
  if (t < n) {
    const uint32_t lane = threadIdx.x;
    const uint32_t i = perm[t];
    const float4 p = leaf[t];
    float4 qlo, qhi;
    sphere_box(p, &qlo, &qhi);
    uint32_t* row = FILL ? neighbors + offsets[i] : nullptr;
    int sp = 0;
    stk[sp++][lane] = 0u;  // the root (n >= 2: node 0 is an inner node)
    uint32_t lpos = 0, lend = 0;
    // One back-edge only: a loop with several back-edges is split by the
    // compiler into nested loops, which serialises the lanes' node and leaf
    // steps.
    for (;;) {
      if (lpos == lend) {
        if (sp == 0) break;
        const uint32_t ref = stk[--sp][lane];
        if (ref & kLeaf) {
          lpos = ref & kLeafPos;
          lend = lpos + ((ref >> 29) & 3u) + 1u;
          // leaf ranges lie inside [0, n) by construction
          if ((uint64_t)lend > (uint64_t)n) lend = (uint32_t)n;
          if (lpos > lend) lpos = lend;
        } else if (ref < (uint32_t)(n - 1)) {
          const BvhNode nd = nodes[ref];
          const bool hl = boxes_overlap(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, qlo, qhi);
          const bool hr = boxes_overlap(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, qlo, qhi);
          // a depth-first walk holds at most one pending sibling per level
          if (hl && sp < (int)stack_rows) stk[sp++][lane] = nd.d.x;
          if (hr && sp < (int)stack_rows) stk[sp++][lane] = nd.d.y;
        }
      }
      // Opaque to the optimiser, which would otherwise thread the node path
      // straight back to the loop header.
      __asm__ volatile("" : "+v"(lpos));
      if (lpos != lend) {
        const uint32_t k = lpos++;
        const uint32_t j = perm[k];
        const float4 q = leaf[k];
        if (j != i) {
          const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          const float d2 = (dx * dx + dy * dy) + dz * dz;
          const float sum = p.w + q.w;
          if (d2 <= sum * sum) {
            if (FILL) row[cnt] = j;
            ++cnt;
          }
        }
      }
    }
    if (!FILL) counts[i] = cnt;
  }
  if (!FILL) {
    unsigned long long c = cnt;
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
    if (threadIdx.x == 0 && c) atomicAdd(total, c);
  }
}

}  // namespace bvk

namespace {

int free_adjacency(bv_adjacency* a) {
  if (!a) return BV_OK;
  if (a->offsets) (void)hipFree(a->offsets);
  if (a->neighbors) (void)hipFree(a->neighbors);
  delete a;
  return BV_OK;
}

// Device temporaries of one build, freed on every exit.
struct Temps {
  void* p[16] = {};
  int k = 0;
  template <class T>
  hipError_t alloc(T** out, size_t bytes) {
    void* q = nullptr;
    if (k == 16) return hipErrorOutOfMemory;
    const hipError_t e = hipMalloc(&q, bytes ? bytes : 4);
    if (e == hipSuccess) p[k++] = q;
    *out = static_cast<T*>(q);
    return e;
  }
  ~Temps() {
    for (int i = 0; i < k; ++i) (void)hipFree(p[i]);
  }
};

#define CECH_HIP(call)                                                                       \
  do {                                                                                       \
    const hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                                  \
      bvh::set_error(std::string(#call) + " failed: " + hipGetErrorString(e_));              \
      return e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation ? BV_ERR_OOM : BV_ERR_HIP; \
    }                                                                                        \
  } while (0)

int build_cech(const float* xyz, const float* radius, int64_t n, hipStream_t stream, bv_adjacency* a) {
  CECH_HIP(hipMalloc((void**)&a->offsets, 4 * (size_t)(n + 1)));
  if (n < 2) {
    // no pair: every row is empty
    CECH_HIP(hipMemsetAsync(a->offsets, 0, 4 * (size_t)(n + 1), stream));
    CECH_HIP(hipStreamSynchronize(stream));
    return BV_OK;
  }
  Temps tmp;
  float *xyz_d = nullptr, *r_d = nullptr;
  float4 *pts = nullptr, *lo = nullptr, *hi = nullptr, *leaf = nullptr;
  bvk::BvhNode* nodes = nullptr;
  uint32_t *perm = nullptr, *counts = nullptr, *unsorted = nullptr;
  unsigned long long* total = nullptr;
  CECH_HIP(tmp.alloc(&xyz_d, 12 * (size_t)n));
  CECH_HIP(tmp.alloc(&r_d, 4 * (size_t)n));
  CECH_HIP(tmp.alloc(&pts, 16 * (size_t)n));
  CECH_HIP(tmp.alloc(&lo, 16 * (size_t)n));
  CECH_HIP(tmp.alloc(&hi, 16 * (size_t)n));
  CECH_HIP(tmp.alloc(&leaf, 16 * (size_t)n));
  CECH_HIP(tmp.alloc(&nodes, sizeof(bvk::BvhNode) * (size_t)(n - 1)));
  CECH_HIP(tmp.alloc(&perm, 4 * (size_t)n));
  CECH_HIP(tmp.alloc(&counts, 4 * (size_t)(n + 1)));
  CECH_HIP(tmp.alloc(&total, 8));
  CECH_HIP(hipMemcpyAsync(xyz_d, xyz, 12 * (size_t)n, hipMemcpyHostToDevice, stream));
  CECH_HIP(hipMemcpyAsync(r_d, radius, 4 * (size_t)n, hipMemcpyHostToDevice, stream));
  const int bs = 256;
  const unsigned grid = (unsigned)((n + bs - 1) / bs);
  hipLaunchKernelGGL(bvk::k_cech_sites, dim3(grid), dim3(bs), 0, stream, n, xyz_d, r_d, pts);
  hipLaunchKernelGGL(bvk::k_cech_boxes, dim3(grid), dim3(bs), 0, stream, n, pts, lo, hi);
  uint32_t depth = 0;
  int rc = bvh::build_lbvh(n, pts, lo, hi, nodes, perm, stream, &depth);
  if (rc != BV_OK) return rc;
  hipLaunchKernelGGL(bvk::k_cech_leaves, dim3(grid), dim3(bs), 0, stream, n, perm, pts, leaf);
  const uint32_t rows = bvk::stack_rows_for(depth);
  const unsigned qgrid = (unsigned)((n + bvk::kSplatBlock - 1) / bvk::kSplatBlock);
  CECH_HIP(hipMemsetAsync(total, 0, 8, stream));
  CECH_HIP(hipMemsetAsync(counts + n, 0, 4, stream));
  hipLaunchKernelGGL(bvk::k_cech_query<false>, dim3(qgrid), dim3(bvk::kSplatBlock), bvk::stack_bytes(rows), stream,
                     n, nodes, perm, leaf, rows, counts, total, (const uint32_t*)nullptr, (uint32_t*)nullptr);
  CECH_HIP(hipGetLastError());
  unsigned long long edges = 0;
  CECH_HIP(hipMemcpyAsync(&edges, total, 8, hipMemcpyDeviceToHost, stream));
  CECH_HIP(hipStreamSynchronize(stream));
  // CSR offsets are u32, and the scan and sort count items in int
  BV_REQUIRE(edges <= 0x7FFFFFFFull, BV_ERR_UNSUPPORTED,
             "Cech graph has " + std::to_string(edges) + " directed edges; at most 2^31 - 1 are supported");
  a->n_edges = edges;
  size_t scan_bytes = 0, sort_bytes = 0;
  CECH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, counts, a->offsets, (int)(n + 1), stream));
  CECH_HIP(hipMalloc((void**)&a->neighbors, 4 * (size_t)std::max<uint64_t>(edges, 1)));
  if (edges == 0) {
    CECH_HIP(hipMemsetAsync(a->offsets, 0, 4 * (size_t)(n + 1), stream));
    CECH_HIP(hipStreamSynchronize(stream));
    return BV_OK;
  }
  CECH_HIP(tmp.alloc(&unsorted, 4 * (size_t)edges));
  int bits = 1;
  while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)n) ++bits;
  CECH_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, sort_bytes, unsorted, a->neighbors, (int)edges,
                                                     (int)n, a->offsets, a->offsets + 1, 0, bits, stream));
  void* work = nullptr;
  CECH_HIP(tmp.alloc(&work, std::max(scan_bytes, sort_bytes)));
  CECH_HIP(hipcub::DeviceScan::ExclusiveSum(work, scan_bytes, counts, a->offsets, (int)(n + 1), stream));
  hipLaunchKernelGGL(bvk::k_cech_query<true>, dim3(qgrid), dim3(bvk::kSplatBlock), bvk::stack_bytes(rows), stream,
                     n, nodes, perm, leaf, rows, (uint32_t*)nullptr, (unsigned long long*)nullptr,
                     (const uint32_t*)a->offsets, unsorted);
  CECH_HIP(hipGetLastError());
  CECH_HIP(hipcub::DeviceSegmentedRadixSort::SortKeys(work, sort_bytes, unsorted, a->neighbors, (int)edges, (int)n,
                                                     a->offsets, a->offsets + 1, 0, bits, stream));
  CECH_HIP(hipStreamSynchronize(stream));
  return BV_OK;
}

}  // namespace

extern "C" {

int bv_compute_cech(const float* xyz, const float* radius, int64_t n_sites, void* stream, bv_adjacency** out) {
  BV_REQUIRE(out, BV_ERR_INVALID, "null argument");
  *out = nullptr;
  BV_REQUIRE(n_sites >= 0 && n_sites < (int64_t)bvk::kLeafPos, BV_ERR_INVALID, "site count out of range");
  BV_REQUIRE(n_sites == 0 || (xyz && radius), BV_ERR_INVALID, "null site arrays");
  bv_adjacency* a = new bv_adjacency();
  a->n = n_sites;
  const int rc = build_cech(xyz, radius, n_sites, static_cast<hipStream_t>(stream), a);
  if (rc != BV_OK) {
    free_adjacency(a);
    return rc;
  }
  *out = a;
  return BV_OK;
}

int bv_adjacency_get_info(const bv_adjacency* adj, int64_t* n_sites, uint64_t* n_edges) {
  BV_REQUIRE(adj, BV_ERR_INVALID, "null adjacency");
  if (n_sites) *n_sites = adj->n;
  if (n_edges) *n_edges = adj->n_edges;
  return BV_OK;
}

int bv_adjacency_read(const bv_adjacency* adj, uint32_t* offsets, uint32_t* neighbors) {
  BV_REQUIRE(adj, BV_ERR_INVALID, "null adjacency");
  if (!adj->host_offsets.empty()) {
    if (offsets) std::memcpy(offsets, adj->host_offsets.data(), 4 * (size_t)(adj->n + 1));
    if (neighbors && adj->n_edges) std::memcpy(neighbors, adj->host_neighbors.data(), 4 * (size_t)adj->n_edges);
    return BV_OK;
  }
  if (offsets) BV_HIP(hipMemcpy(offsets, adj->offsets, 4 * (size_t)(adj->n + 1), hipMemcpyDeviceToHost));
  if (neighbors && adj->n_edges)
    BV_HIP(hipMemcpy(neighbors, adj->neighbors, 4 * (size_t)adj->n_edges, hipMemcpyDeviceToHost));
  return BV_OK;
}

int bv_adjacency_destroy(bv_adjacency* adj) { return free_adjacency(adj); }

}  // extern "C"
