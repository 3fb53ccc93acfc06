// Compaction of recorded path rows into one dense record stream.
//
// Reference semantics: PathCompactor::dispatch (blade-volume/src/gpu/
// path_compact.rs:63-89) running compact_paths (shaders/path_compact.wgsl:
// 24-50): every path's live prefix (status & 0x7fffffff steps of its
// max_steps-wide row) is appended to the output at `count`, one record per
// step: the dense slot it came from (pixel * max_steps + step), its cell, its
// pixel and an active weight of 1.  The reference claims each path's block
// with one atomic, so the order of blocks follows scheduling; here the blocks
// follow path order (an exclusive scan), one of the orders the reference can
// produce, and the same on every run.
//
// Design on gfx950 (HBM-bound byte work: 4 B of status per path, 4 B read and
// 16 B written per record):
//   * k_compact_sums: 256 paths per 256-lane block, a block sum of their live
//     steps;
//   * k_compact_offsets: one block scans the block sums into the output start
//     of every block (after the `count` already there) and stores the new
//     count;
//   * k_compact_write: each block rescans its 256 step counts in LDS, and its
//     lanes stride over the block's contiguous output range in aligned quads
//     of records, finding a quad's first path by a binary search of the LDS
//     prefix and its other records by stepping on, so every stream is written
//     with consecutive lanes on consecutive 16-byte words.
#include <hip/hip_runtime.h>

#include <string>

#include "bv_host.h"

namespace bvk {
namespace {

constexpr int kCompactThreads = 256;
// One path per lane.  For one 1024x768 camera's 128-step rows (63M records)
// the write pass measured 417 us; 4 paths per lane (a quarter of the waves to
// hide the search's LDS latency) 603 us; one-wave blocks of 64 paths 447 us
// (and a 101 us scan of their block sums).
constexpr int kPathsPerBlock = 256;
constexpr int kPerLane = kPathsPerBlock / kCompactThreads;

__device__ __forceinline__ uint32_t live_steps(const uint32_t* __restrict__ status, uint64_t p, uint64_t n,
                                               uint32_t max_steps) {
  if (p >= n) return 0u;
  const uint32_t s = status[p] & 0x7fffffffu;
  return s < max_steps ? s : max_steps;  // a row never holds more than max_steps records
}

// Exclusive prefix of v over the block's 256 lanes; *total receives the sum.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wave_sums[kCompactThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wave_sums[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kCompactThreads / 64; ++w) {
    const uint32_t s = wave_sums[w];
    before += w < wave ? s : 0u;
    all += s;
  }
  __syncthreads();  // wave_sums is reused by the next call
  *total = all;
  return before + x - v;
}

__global__ void __launch_bounds__(kCompactThreads)
    k_compact_sums(const uint32_t* __restrict__ status, uint64_t n, uint32_t max_steps, uint32_t* block_sums) {
  const uint64_t first = (uint64_t)blockIdx.x * kPathsPerBlock;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kPerLane; ++k)
    s += live_steps(status, first + k * kCompactThreads + threadIdx.x, n, max_steps);
  uint32_t total;
  (void)block_exclusive(s, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// One block: block_start[b] = *count + sum of block_sums[0 .. b); *count += the total.
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_offsets(const uint32_t* __restrict__ block_sums, uint32_t n_blocks, uint32_t* block_start,
                      uint32_t* count) {
  // 64-bit so that an append past 2^32 records saturates instead of wrapping
  // (a saturated start lies beyond every capacity, so nothing is written)
  uint64_t running = *count;
  for (uint32_t b0 = 0; b0 < n_blocks; b0 += kCompactThreads) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t v = b < n_blocks ? block_sums[b] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive(v, &total);
    if (b < n_blocks) block_start[b] = (uint32_t)min(running + ex, (uint64_t)0xffffffffu);
    running += total;
  }
  if (threadIdx.x == 0) *count = (uint32_t)min(running, (uint64_t)0xffffffffu);
}

typedef uint32_t U4 __attribute__((ext_vector_type(4)));
typedef float F4 __attribute__((ext_vector_type(4)));

// Last path of the block whose start is <= pos (empty paths share their
// successor's start, and the search passes over them).
__device__ __forceinline__ uint32_t path_of(const uint32_t* start, uint32_t pos) {
  uint32_t lo = 0, hi = kPathsPerBlock - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (start[mid] <= pos) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Each lane writes aligned quads of the output: records 4q .. 4q + 3 of the
// stream (those inside the block's range), with one path search per quad
// (the quad's later records advance from it) and, when the streams are 16 B
// aligned and the quad is whole, one 16-byte store per stream.
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_write(const uint32_t* __restrict__ status, const uint32_t* __restrict__ cells, uint64_t n,
                    uint32_t max_steps, const uint32_t* __restrict__ block_start, uint64_t capacity, bool vec,
                    uint32_t* dense_slots, uint32_t* out_cells, uint32_t* pixel_indices, float* active) {
  __shared__ uint32_t start[kPathsPerBlock];
  const uint64_t first = (uint64_t)blockIdx.x * kPathsPerBlock;
  // lane t owns paths kPerLane t .. kPerLane (t + 1) - 1 of the block, so its
  // prefix is one scan
  uint32_t c[kPerLane], s = 0;
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    c[k] = live_steps(status, first + kPerLane * threadIdx.x + k, n, max_steps);
    s += c[k];
  }
  uint32_t total;
  uint32_t at = block_exclusive(s, &total);
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    start[kPerLane * threadIdx.x + k] = at;
    at += c[k];
  }
  __syncthreads();
  const uint64_t out0 = block_start[blockIdx.x];
  const uint64_t q_end = (out0 + total + 3) >> 2;
  for (uint64_t q = (out0 >> 2) + threadIdx.x; q < q_end; q += kCompactThreads) {
    const int64_t lp0 = (int64_t)(q << 2) - (int64_t)out0;  // block position of the quad's first record
    uint32_t lo = path_of(start, lp0 < 0 ? 0u : (uint32_t)lp0);
    uint32_t slot[4], cell[4], pix[4];
    bool in[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t lp = lp0 + e;
      in[e] = lp >= 0 && lp < (int64_t)total && (q << 2) + e < capacity;
      slot[e] = 0;
      cell[e] = 0;
      pix[e] = 0;
      if (lp >= 0 && lp < (int64_t)total) {
        while (lo + 1 < (uint32_t)kPathsPerBlock && start[lo + 1] <= (uint32_t)lp) ++lo;
        const uint64_t pixel = first + lo;
        const uint64_t sl = pixel * max_steps + ((uint32_t)lp - start[lo]);
        slot[e] = (uint32_t)sl;
        pix[e] = (uint32_t)pixel;
        if (in[e]) cell[e] = cells[sl];
      }
    }
    const uint64_t o = q << 2;
    if (vec && in[0] && in[1] && in[2] && in[3]) {
      // streaming (nontemporal) stores: the stream is not read back by this
      // kernel; measured 0.366 ms per call against 0.418 with plain stores
      __builtin_nontemporal_store(U4{slot[0], slot[1], slot[2], slot[3]}, reinterpret_cast<U4*>(dense_slots) + q);
      __builtin_nontemporal_store(U4{cell[0], cell[1], cell[2], cell[3]}, reinterpret_cast<U4*>(out_cells) + q);
      __builtin_nontemporal_store(U4{pix[0], pix[1], pix[2], pix[3]}, reinterpret_cast<U4*>(pixel_indices) + q);
      __builtin_nontemporal_store(F4{1.0f, 1.0f, 1.0f, 1.0f}, reinterpret_cast<F4*>(active) + q);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (in[e]) {
          dense_slots[o + e] = slot[e];
          out_cells[o + e] = cell[e];
          pixel_indices[o + e] = pix[e];
          active[o + e] = 1.0f;
        }
      }
    }
  }
}

}  // namespace
}  // namespace bvk

extern "C" {

int bv_compact_paths(const uint32_t* path_status, const uint32_t* cells, uint32_t num_pixels, uint32_t max_steps,
                     uint32_t* count, uint32_t* dense_slots, uint32_t* compact_cells, uint32_t* pixel_indices,
                     float* active, uint64_t capacity, void* stream) {
  BV_REQUIRE(count, BV_ERR_INVALID, "null count");
  BV_REQUIRE((uint64_t)num_pixels * max_steps <= 0xffffffffull, BV_ERR_INVALID,
             "num_pixels * max_steps must fit the 32-bit dense slot index");
  if (num_pixels == 0 || max_steps == 0) return BV_OK;
  BV_REQUIRE(path_status && cells, BV_ERR_INVALID, "null path rows");
  BV_REQUIRE(capacity < 0xffffffffull, BV_ERR_INVALID, "capacity must be below 2^32 - 1 records");
  BV_REQUIRE(capacity == 0 || (dense_slots && compact_cells && pixel_indices && active), BV_ERR_INVALID,
             "null output stream");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint32_t n_blocks = (num_pixels + bvk::kPathsPerBlock - 1) / bvk::kPathsPerBlock;
  uint32_t* tmp = nullptr;
  BV_HIP(hipMallocAsync(reinterpret_cast<void**>(&tmp), 8 * (size_t)n_blocks, s));
  uint32_t* sums = tmp;
  uint32_t* starts = tmp + n_blocks;
  hipLaunchKernelGGL(bvk::k_compact_sums, dim3(n_blocks), dim3(bvk::kCompactThreads), 0, s, path_status,
                     (uint64_t)num_pixels, max_steps, sums);
  hipLaunchKernelGGL(bvk::k_compact_offsets, dim3(1), dim3(bvk::kCompactThreads), 0, s, sums, n_blocks, starts,
                     count);
  // 16-byte stores need every stream 16-byte aligned
  const bool vec = ((reinterpret_cast<uintptr_t>(dense_slots) | reinterpret_cast<uintptr_t>(compact_cells) |
                     reinterpret_cast<uintptr_t>(pixel_indices) | reinterpret_cast<uintptr_t>(active)) &
                    15u) == 0u;
  hipLaunchKernelGGL(bvk::k_compact_write, dim3(n_blocks), dim3(bvk::kCompactThreads), 0, s, path_status, cells,
                     (uint64_t)num_pixels, max_steps, starts, capacity, vec, dense_slots, compact_cells,
                     pixel_indices, active);
  const hipError_t launch = hipGetLastError();
  BV_HIP(hipFreeAsync(tmp, s));
  BV_HIP(launch);
  return BV_OK;
}

}  // extern "C"
