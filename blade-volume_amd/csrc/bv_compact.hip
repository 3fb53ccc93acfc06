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
//     of every block (after the `count` already there), stores the new count
//     and the call's output range;
//   * k_compact_starts: each block rescans its 256 step counts and writes
//     every path's output start, and for every 256-record tile of the output
//     the path its first record belongs to;
//   * k_compact_tiles: a resident grid whose waves take consecutive tiles of
//     the output, so the device writes each stream as one advancing front
//     (with a block per path range, thousands of separate write fronts ran
//     at 3.4 TB/s, against about 5 TB/s for the same bytes written as a
//     front).  Lane l writes the tile's records 4l .. 4l + 3, finding their
//     paths by stepping on from the tile's first path.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// One block: block_start[b] = *count + sum of block_sums[0 .. b); *count += the total;
// bounds = (count before, count after) of this call.
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_offsets(const uint32_t* __restrict__ block_sums, uint32_t n_blocks, uint32_t* block_start,
                      uint32_t* count, uint32_t* bounds) {
  // 64-bit so that an append past 2^32 records saturates instead of wrapping
  // (a saturated start lies beyond every capacity, so nothing is written)
  const uint64_t before = *count;
  uint64_t running = before;
  for (uint32_t b0 = 0; b0 < n_blocks; b0 += kCompactThreads) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t v = b < n_blocks ? block_sums[b] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive(v, &total);
    if (b < n_blocks) block_start[b] = (uint32_t)min(running + ex, (uint64_t)0xffffffffu);
    running += total;
  }
  if (threadIdx.x == 0) {
    *count = (uint32_t)min(running, (uint64_t)0xffffffffu);
    bounds[0] = (uint32_t)before;
    bounds[1] = (uint32_t)min(running, (uint64_t)0xffffffffu);
  }
}

typedef uint32_t U4 __attribute__((ext_vector_type(4)));
typedef float F4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kTileRecords = 256;  // one wave: 64 lanes x 4 records

// Every path's output start, and tile_first[t] = the path holding the first
// record of output tile t0 + t (t0 = the tile of the call's first record; a
// tile starts at the later of its aligned position and the call's start).
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_starts(const uint32_t* __restrict__ status, uint64_t n, uint32_t max_steps,
                     const uint32_t* __restrict__ block_start, const uint32_t* __restrict__ bounds,
                     uint32_t* path_start, uint32_t* tile_first) {
  const uint64_t first = (uint64_t)blockIdx.x * kPathsPerBlock;
  uint32_t c[kPerLane], s = 0;
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    c[k] = live_steps(status, first + kPerLane * threadIdx.x + k, n, max_steps);
    s += c[k];
  }
  uint32_t total;
  uint64_t at = (uint64_t)block_start[blockIdx.x] + block_exclusive(s, &total);
  const uint64_t c0 = bounds[0];
  const uint64_t t0 = c0 / kTileRecords;
#pragma unroll
  for (int k = 0; k < kPerLane; ++k) {
    const uint64_t p = first + kPerLane * threadIdx.x + k;
    if (p < n) {
      path_start[p] = (uint32_t)min(at, (uint64_t)0xffffffffu);
      if (c[k] > 0) {
        for (uint64_t t = at / kTileRecords; t <= (at + c[k] - 1) / kTileRecords; ++t) {
          const uint64_t x = max(t * kTileRecords, c0);
          if (x >= at) tile_first[t - t0] = (uint32_t)p;
        }
      }
    }
    at += c[k];
  }
}

// A wave per tile: lane i loads the start of path p0 + i (p0 = the path of the
// tile's first record; one coalesced load) into a 64-entry LDS row, each path
// marks the first quad of the tile whose first record it covers, and a prefix
// maximum over the lanes gives each lane's quad its first path; the quad's
// later records step on through the row.  A tile whose records run past the
// 64 paths held (paths of under four records on average) steps on through
// memory.
__global__ void __launch_bounds__(kCompactThreads)
    k_compact_tiles(const uint32_t* __restrict__ cells, uint64_t n, uint32_t max_steps,
                    const uint32_t* __restrict__ path_start, const uint32_t* __restrict__ tile_first,
                    const uint32_t* __restrict__ bounds, uint64_t capacity, bool vec, uint32_t* dense_slots,
                    uint32_t* out_cells, uint32_t* pixel_indices, float* active) {
  __shared__ int first_path[kCompactThreads];         // 64 entries per wave
  __shared__ uint32_t starts_row[kCompactThreads];   // 64 path starts per wave
  const uint64_t c0 = bounds[0];
  const uint64_t c_end = min((uint64_t)bounds[1], capacity);  // records past the capacity are not written
  if (c_end <= c0) return;
  const uint64_t t0 = c0 / kTileRecords, t_end = (c_end - 1) / kTileRecords + 1;
  const uint32_t lane = threadIdx.x & 63u;
  int* row = first_path + (threadIdx.x & ~63u);
  uint32_t* srow = starts_row + (threadIdx.x & ~63u);
  const uint64_t waves = (uint64_t)gridDim.x * (kCompactThreads / 64);
  for (uint64_t t = t0 + (uint64_t)blockIdx.x * (kCompactThreads / 64) + (threadIdx.x >> 6); t < t_end;
       t += waves) {
    const uint64_t tile0 = t * kTileRecords;
    const uint32_t p0 = tile_first[t - t0];
    const bool held = p0 + lane < n;
    const uint32_t my_start = held ? path_start[p0 + lane] : 0xffffffffu;
    row[lane] = -1;
    srow[lane] = my_start;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // the first quad whose first record this path covers (path 0 of the row
    // covers the tile's first quad)
    if (held) {
      const uint64_t q = lane == 0 || my_start <= tile0 ? 0 : (my_start - tile0 + 3) / 4;
      if (q < 64) atomicMax(&row[q], (int)lane);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int pl = row[lane];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(pl, off, 64);
      pl = (int)lane >= off && o > pl ? o : pl;
    }
    const uint64_t r0 = tile0 + 4u * lane;  // this lane's quad: records r0 .. r0 + 3
    uint32_t slot[4], cell[4], pix[4];
    bool in[4];
    uint32_t p = p0 + (uint32_t)(pl < 0 ? 0 : pl);
    uint32_t ps = srow[pl < 0 ? 0 : pl];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint64_t r = r0 + e;
      in[e] = r >= c0 && r < c_end;
      // step on: the next path's start, from the row or (past it) memory
      for (;;) {
        const uint32_t k = p - p0 + 1;
        const uint32_t nx = k < 64 ? srow[k] : (p + 1 < n ? path_start[p + 1] : 0xffffffffu);
        if (!(in[e] && p + 1 < n && nx <= (uint32_t)r)) break;
        ++p;
        ps = nx;
      }
      slot[e] = 0;
      cell[e] = 0;
      pix[e] = 0;
      if (in[e]) {
        const uint64_t sl = (uint64_t)p * max_steps + ((uint32_t)r - ps);
        slot[e] = (uint32_t)sl;
        pix[e] = p;
        cell[e] = cells[sl];
      }
    }
    if (vec && in[0] && in[1] && in[2] && in[3]) {
      // streaming (nontemporal) stores: the stream is not read back here
      const uint64_t q = r0 >> 2;
      __builtin_nontemporal_store(U4{slot[0], slot[1], slot[2], slot[3]}, reinterpret_cast<U4*>(dense_slots) + q);
      __builtin_nontemporal_store(U4{cell[0], cell[1], cell[2], cell[3]}, reinterpret_cast<U4*>(out_cells) + q);
      __builtin_nontemporal_store(U4{pix[0], pix[1], pix[2], pix[3]}, reinterpret_cast<U4*>(pixel_indices) + q);
      __builtin_nontemporal_store(F4{1.0f, 1.0f, 1.0f, 1.0f}, reinterpret_cast<F4*>(active) + q);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (in[e]) {
          dense_slots[r0 + e] = slot[e];
          out_cells[r0 + e] = cell[e];
          pixel_indices[r0 + e] = pix[e];
          active[r0 + e] = 1.0f;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
  // tiles of the call's output: at most every record of every row, plus the
  // partial tiles at both ends
  const uint64_t n_tiles = (uint64_t)num_pixels * max_steps / bvk::kTileRecords + 2;
  uint32_t* tmp = nullptr;
  const size_t words = 2 * (size_t)n_blocks + 2 + (size_t)num_pixels + (size_t)n_tiles;
  BV_HIP(hipMallocAsync(reinterpret_cast<void**>(&tmp), 4 * words, s));
  uint32_t* sums = tmp;
  uint32_t* starts = sums + n_blocks;
  uint32_t* bounds = starts + n_blocks;
  uint32_t* path_start = bounds + 2;
  uint32_t* tile_first = path_start + num_pixels;
  hipLaunchKernelGGL(bvk::k_compact_sums, dim3(n_blocks), dim3(bvk::kCompactThreads), 0, s, path_status,
                     (uint64_t)num_pixels, max_steps, sums);
  hipLaunchKernelGGL(bvk::k_compact_offsets, dim3(1), dim3(bvk::kCompactThreads), 0, s, sums, n_blocks, starts,
                     count, bounds);
  hipLaunchKernelGGL(bvk::k_compact_starts, dim3(n_blocks), dim3(bvk::kCompactThreads), 0, s, path_status,
                     (uint64_t)num_pixels, max_steps, starts, bounds, path_start, tile_first);
  // 16-byte stores need every stream 16-byte aligned
  const bool vec = ((reinterpret_cast<uintptr_t>(dense_slots) | reinterpret_cast<uintptr_t>(compact_cells) |
                     reinterpret_cast<uintptr_t>(pixel_indices) | reinterpret_cast<uintptr_t>(active)) &
                    15u) == 0u;
  hipError_t launch = hipGetLastError();
  if (launch == hipSuccess && capacity > 0) {
    int dev = 0, cus = 0, per_cu = 0;
    BV_HIP(hipGetDevice(&dev));
    BV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bvk::k_compact_tiles, bvk::kCompactThreads, 0));
    const uint64_t need = (n_tiles + bvk::kCompactThreads / 64 - 1) / (bvk::kCompactThreads / 64);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)std::max(1, cus * per_cu)));
    hipLaunchKernelGGL(bvk::k_compact_tiles, dim3(grid), dim3(bvk::kCompactThreads), 0, s, cells,
                       (uint64_t)num_pixels, max_steps, path_start, tile_first, bounds, capacity, vec, dense_slots,
                       compact_cells, pixel_indices, active);
    launch = hipGetLastError();
  }
  BV_HIP(hipFreeAsync(tmp, s));
  BV_HIP(launch);
  return BV_OK;
}

}  // extern "C"
