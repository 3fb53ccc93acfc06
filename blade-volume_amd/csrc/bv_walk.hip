// Cooperative walk: G lanes of a wave64 share one ray.
//
// Per step a ray group reads its current cell's row of edge records (32 B per
// face, contiguous, one coalesced request per group) together with the cell's
// colour row, each lane tests its faces with the plane test of
// trace.rs:186-202, and the group takes the lexicographic (t, face index)
// minimum with DPP lane exchanges — the sequential walk's "first face wins on
// ties".  The SH colour is spread over the group too (one component per lane).  The winning record already holds the next cell's position, row
// bounds and radius, so the next step needs no further dependent load: one
// memory round trip per visited cell.  Ray state is replicated in every lane
// of the group, so control flow is uniform within a group and only lane 0
// writes outputs.
//
// Every plane test forms o = p_j - p_i from the two FP32 positions at run time
// and divides with IEEE rounding, exactly as the reference does, so the cell
// sequences equal the oracle's bit for bit for every G.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>

#include "bv_host.h"

namespace {

using namespace bvk;

struct Cell {
  uint32_t id, begin, end;
  V3 p;
  float r;
};

template <int G>
__device__ __forceinline__ uint32_t glane() {
  return threadIdx.x & (G - 1);
}

// Lane exchange inside a ray group.  Groups are aligned runs of G lanes; for
// G <= 16 a group lies in one DPP row, so the butterfly uses DPP row
// permutations (quad xor 1, quad xor 2, half-row mirror, row mirror), each a
// plain VALU operand modifier; the 32-lane step crosses rows with a permute.
// Every combine below is commutative, so all lanes of a group end with
// bit-identical values and the replicated ray state stays identical.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // lane i <-> 7 - i within each half row
constexpr int kDppMirror = 0x140;     // lane i <-> 15 - i within the row

// Every permutation here reads a lane of the same row, so no lane takes the
// "old" operand: mov_dpp leaves it undefined, which spares the register copy
// that initialising it costs.
template <int CTRL>
__device__ __forceinline__ float xchg(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ uint32_t xchg(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

template <int CTRL>
__device__ __forceinline__ void min_step(float& t, uint32_t& j) {
  float ot;
  uint32_t oj;
  if (CTRL == 0) {
    ot = __shfl_xor(t, 16, 64);
    oj = __shfl_xor(j, 16, 64);
  } else {
    ot = xchg<CTRL>(t);
    oj = xchg<CTRL>(j);
  }
  const bool take = ot < t || (ot == t && oj < j);
  t = take ? ot : t;
  j = take ? oj : j;
}

// Lexicographic (t, j) minimum over the G lanes of the ray group.
template <int G>
__device__ __forceinline__ void group_min(float& t, uint32_t& j) {
  if (G >= 2) min_step<kDppXor1>(t, j);
  if (G >= 4) min_step<kDppXor2>(t, j);
  if (G >= 8) min_step<kDppHalfMirror>(t, j);
  if (G >= 16) min_step<kDppMirror>(t, j);
  if (G >= 32) min_step<0>(t, j);
}

template <int CTRL>
__device__ __forceinline__ float4 add_step(float4 v) {
  if (CTRL == 0)
    return make_float4(v.x + __shfl_xor(v.x, 16, 64), v.y + __shfl_xor(v.y, 16, 64), v.z + __shfl_xor(v.z, 16, 64),
                       v.w + __shfl_xor(v.w, 16, 64));
  return make_float4(v.x + xchg<CTRL>(v.x), v.y + xchg<CTRL>(v.y), v.z + xchg<CTRL>(v.z), v.w + xchg<CTRL>(v.w));
}

template <int G>
__device__ __forceinline__ float4 group_sum(float4 v) {
  if (G >= 2) v = add_step<kDppXor1>(v);
  if (G >= 4) v = add_step<kDppXor2>(v);
  if (G >= 8) v = add_step<kDppHalfMirror>(v);
  if (G >= 16) v = add_step<kDppMirror>(v);
  if (G >= 32) v = add_step<0>(v);
  return v;
}

// Loads of one cell that are issued a step ahead: the first G*U edge records of
// its row (lane l holds faces begin + l + u G) and the lane's colour-row
// components (or component 0 only, for its density).
template <int U, int KM>
struct Prefetch {
  float4 a[U], b[U];
  float4 col[KM];
};

// Every lane issues every load (a face index past the row reads the row's
// first record, or the padding record after the last edge; a component past K
// reads component K-1 and is masked), so the number of loads in flight is the
// same on every path and the compiler's wait counts stay exact.
template <int G, int U, int K, int KM, bool COLOUR>
__device__ __forceinline__ void prefetch(const SceneView& s, const Cell& c, Prefetch<U, KM>* f) {
  const uint32_t lane = glane<G>();
  const float4* er = reinterpret_cast<const float4*>(s.erec);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t j = c.begin + lane + u * G;
    const uint32_t jj = j < c.end ? j : c.begin;
    f->a[u] = er[2 * (size_t)jj];
    f->b[u] = er[2 * (size_t)jj + 1];
  }
  const float4* row = reinterpret_cast<const float4*>(s.shd) + (size_t)c.id * K;
  if (COLOUR) {
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      const uint32_t k = lane + m * G;
      const float4 v = row[k < (uint32_t)K ? k : (uint32_t)(K - 1)];
      f->col[m] = (KM * G == K || k < (uint32_t)K) ? v : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  } else {
    f->col[0] = row[0];
  }
}

// One pass of plane tests (trace.rs:186-202) over this lane's faces.
template <int G, int U, bool BOUNDED, bool CARRY>
__device__ __forceinline__ void test_pass(const float4 (&a)[U], const float4 (&b)[U], uint32_t base, const Cell& c,
                                          V3 ro, V3 d, float t0, float* bt, uint32_t* bj, float4* wa, float4* wb) {
  const uint32_t lane = glane<G>();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t j = base + lane + u * G;
    if (CARRY) {
      // the candidate's record goes with it (assigned under the condition: a
      // select would become a select of addresses and put the record arrays
      // in scratch)
      if (j < c.end) {
        const V3 o = sub(mk(a[u].x, a[u].y, a[u].z), c.p);
        const V3 fo = add(c.p, scl(BOUNDED ? b[u].w : 0.5f, o));
        const float dp = dot(o, d);
        if (dp > 0.0f) {
          const float t = dot(sub(fo, ro), o) / dp;
          if (isfinite(t) && t > t0 && t < *bt) {
            *bt = t;
            *bj = j;
            *wa = a[u];
            *wb = b[u];
          }
        }
      }
    } else {
      // branch-free: selects instead of exec-mask branches and register copies
      const V3 o = sub(mk(a[u].x, a[u].y, a[u].z), c.p);
      const V3 fo = add(c.p, scl(BOUNDED ? b[u].w : 0.5f, o));
      const float dp = dot(o, d);
      const float t = dot(sub(fo, ro), o) / dp;
      const bool take = j < c.end && dp > 0.0f && isfinite(t) && t > t0 && t < *bt;
      *bt = take ? t : *bt;
      *bj = take ? j : *bj;
    }
  }
}

// Exit face of cell c: sets *t1 (depth if none) and, when a face exists, *next.
// The first pass uses the prefetched records; longer rows load the rest here.
template <int G, int U, bool BOUNDED, bool RELOAD = true, int KM>
__device__ __forceinline__ bool coop_exit(const SceneView& s, const Cell& c, const Prefetch<U, KM>& f, V3 ro, V3 d,
                                          float t0, float depth, float* t1, Cell* next) {
  const uint32_t lane = glane<G>();
  float bt = depth;
  uint32_t bj = BV_NO_CELL;
  // RELOAD: each lane tracks (t, edge) only, and the group reads the winner's
  // record again after the argmin (one address, a line it loaded a moment
  // ago), which spares the register copies of carrying candidate records.
  // That read waits, in order, behind every store and atomic the wave still
  // has in flight, so walks that store per step (the path cache, recorder
  // rows) or add gradients carry the records and shuffle the winner's out.
  float4 wa = make_float4(0.f, 0.f, 0.f, 0.f), wb = wa;
  test_pass<G, U, BOUNDED, !RELOAD>(f.a, f.b, c.begin, c, ro, d, t0, &bt, &bj, &wa, &wb);
  const float4* er = reinterpret_cast<const float4*>(s.erec);
  for (uint32_t base = c.begin + G * U; base < c.end; base += G * U) {
    float4 a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t j = base + lane + u * G;
      const uint32_t jj = j < c.end ? j : c.begin;
      a[u] = er[2 * (size_t)jj];
      b[u] = er[2 * (size_t)jj + 1];
    }
    test_pass<G, U, BOUNDED, !RELOAD>(a, b, base, c, ro, d, t0, &bt, &bj, &wa, &wb);
  }
  float t = bt;
  uint32_t j = bj;
  if (G > 1) group_min<G>(t, j);
  *t1 = t;
  if (j == BV_NO_CELL) return false;
  if (RELOAD) {
    wa = er[2 * (size_t)j];
    wb = er[2 * (size_t)j + 1];
  } else if (G > 1) {
    const int src = (int)((threadIdx.x & 63u) & ~(uint32_t)(G - 1)) + (int)((j - c.begin) & (G - 1));
    wa.x = __shfl(wa.x, src, 64);
    wa.y = __shfl(wa.y, src, 64);
    wa.z = __shfl(wa.z, src, 64);
    wa.w = __shfl(wa.w, src, 64);
    wb.x = __shfl(wb.x, src, 64);
    wb.y = __shfl(wb.y, src, 64);
    wb.z = __shfl(wb.z, src, 64);
  }
  next->id = __float_as_uint(wa.w);
  next->p = mk(wa.x, wa.y, wa.z);
  next->begin = __float_as_uint(wb.x);
  next->end = __float_as_uint(wb.y);
  next->r = wb.z;
  return true;
}

__device__ __forceinline__ Cell start_cell(const SceneView& s, uint32_t id) {
  const CellRec c = s.cells[id];
  return Cell{id, c.begin, c.end, mk(c.x, c.y, c.z), c.radius};
}

// Colour of a cell, spread over the group: lane l owns SH components
// l, l + G, ... of the component-major row (float4 (r, g, b, .) per component,
// the density in component 0's fourth slot).  Each lane multiplies its
// coefficients by its basis values C_k Y_k(d) (fixed per ray) and the group
// sums the partial products; the fourth channel sums to the density exactly,
// since every other component stores zero there.
template <int G, int DEG>
struct Colour {
  static constexpr int K = (DEG + 1) * (DEG + 1);
  static constexpr int KM = (K + G - 1) / G;
  float basis[KM];

  __device__ __forceinline__ void init(V3 dir) {
    float all[16];
    sh_basis(K, dir, all);
    const uint32_t lane = threadIdx.x & (G - 1);
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      const uint32_t k = lane + m * G;
      float v = 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q) v = (k == (uint32_t)q) ? all[q] : v;
      basis[m] = v;
    }
  }
  // (r, g, b) = unclamped 0.5 + SH colour, w = density, from the lane's
  // prefetched components.
  __device__ __forceinline__ float4 eval(const float4 (&c)[KM]) const {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      acc.x += c[m].x * basis[m];
      acc.y += c[m].y * basis[m];
      acc.z += c[m].z * basis[m];
      acc.w += c[m].w;
    }
    acc = group_sum<G>(acc);
    return make_float4(0.5f + acc.x, 0.5f + acc.y, 0.5f + acc.z, acc.w);
  }
};


__device__ __forceinline__ float pick3(V3 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Gradient scatter of one integrated segment: 3K SH values and ten scalar
// slots (density; current, next and previous xyz) form N items, striped over
// the group's lanes in R rounds (item = lane + r G).  Each round is one atomic
// instruction on every lane; the per-lane item of a round is fixed for the
// ray, so its SH basis factor is precomputed.  Leftover lanes of the last
// round add zero to distinct SH entries of the same cell, so no two lanes of
// one instruction share an address.
template <int G, int K>
struct Scatter {
  static constexpr int N = 3 * K + 10;
  static constexpr int R = (N + G - 1) / G;
  float bval[R];

  __device__ __forceinline__ void init(V3 dir) {
    float all[16];
    sh_basis(K, dir, all);
    const uint32_t lane = threadIdx.x & (G - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t item = lane + r * G;
      const uint32_t k = item / 3u;
      float v = 0.f;
#pragma unroll
      for (int q = 0; q < K; ++q) v = (k == (uint32_t)q) ? all[q] : v;
      bval[r] = item < 3u * K ? v : 0.f;
    }
  }

  // One replayed step with its cell's position gradient already summed over
  // the faces it shares with the neighbouring steps (gxyz); the chunk's first
  // and last steps also add the face terms of their neighbours outside the
  // chunk (gprev to prev when bprev, gnext to next when bnext).  One cell's
  // SH row, density and position per step instead of three positions, and no
  // padding lanes: the fourth instruction touches two lines, not five.
  __device__ __forceinline__ void add_merged(const bvh::GradPtrs& gr, uint32_t cell, V3 gc, float d_sigma, V3 gxyz,
                                             bool bprev, uint32_t prev, V3 gprev, bool bnext, uint32_t next,
                                             V3 gnext) const {
    const uint32_t lane = threadIdx.x & (G - 1);
    float* dsh = gr.d_sh + (size_t)cell * 3 * K;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t item = lane + r * G;
      float* addr = dsh + item;
      float val = 0.f;
      bool live = true;
      if (item < 3u * K) {
        val = pick3(gc, item % 3u) * bval[r];
      } else {
        const uint32_t q = item - 3u * K;
        const uint32_t part = q == 0 ? 0u : (q < 4 ? 1u : (q < 7 ? 2u : 3u));
        const uint32_t comp = q == 0 ? 0u : (q - 1u) % 3u;
        live = item < (uint32_t)N &&
               (part <= 1 || (part == 2 && bprev) || (part == 3 && bnext));
        const uint32_t at = part == 2 ? prev : (part == 3 ? next : cell);
        const V3 gv = part == 1 ? gxyz : (part == 2 ? gprev : gnext);
        if (gr.packed) {
          float* slot = reinterpret_cast<float*>(gr.packed + at);
          addr = part == 0 ? slot + 3 : slot + comp;
        } else {
          addr = part == 0 ? gr.d_density + cell : gr.d_xyz + 3 * (size_t)at + comp;
        }
        val = part == 0 ? d_sigma : pick3(gv, comp);
      }
      // lanes with nothing to add issue no atomic: lanes of one instruction on
      // one address are serialised into separate memory requests
      if (live) atomicAdd(addr, val);
    }
  }
};

// The first kHotSteps steps of the rays a block replays, summed per cell in
// LDS and added to the gradient tables once, when the block ends.  Rays that
// share a start cell (a camera's rays) cross the same few cells in their first
// steps: hundreds of thousands of rays adding to one cell serialise in the
// memory atomics (on 786K rays of one camera, the memory atomics of the first
// step alone took half the call).  Open addressing over 64 slots, two probes;
// a slot is claimed by the first cell that lands on it and held for the
// launch, and a step whose probes are held by other cells adds to memory as
// before.
constexpr uint32_t kHotSteps = 4;

template <int K>
struct HotCells {
  static constexpr int kSlots = 64;
  static constexpr int kRow = 3 * K + 4;  // SH row, then x, y, z, density
  uint32_t key[kSlots];
  float val[kSlots][kRow];

  __device__ __forceinline__ void clear() {
    for (uint32_t i = threadIdx.x; i < (uint32_t)(kSlots * kRow); i += blockDim.x) (&val[0][0])[i] = 0.f;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kSlots; i += blockDim.x) key[i] = BV_NO_CELL;
  }

  // The cell's slot (claimed when free), or -1.
  __device__ __forceinline__ int find(uint32_t cell) {
    const uint32_t h = (cell * 0x9E3779B1u) >> 26;
#pragma unroll
    for (uint32_t p = 0; p < 2; ++p) {
      const uint32_t at = (h + p) & (kSlots - 1);
      const uint32_t old = atomicCAS(&key[at], BV_NO_CELL, cell);
      if (old == BV_NO_CELL || old == cell) return (int)at;
    }
    return -1;
  }

  // Every held slot's sums into the tables (after a block barrier).
  __device__ __forceinline__ void flush(const bvh::GradPtrs& gr) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)(kSlots * kRow); i += blockDim.x) {
      const uint32_t at = i / kRow, item = i % kRow;
      const uint32_t cell = key[at];
      const float v = val[at][item];
      if (cell == BV_NO_CELL || v == 0.f) continue;
      float* addr;
      if (item < 3u * K) {
        addr = gr.d_sh + (size_t)cell * 3 * K + item;
      } else {
        const uint32_t q = item - 3u * K;  // x, y, z, density
        if (gr.packed)
          addr = reinterpret_cast<float*>(gr.packed + cell) + q;
        else
          addr = q == 3 ? gr.d_density + cell : gr.d_xyz + 3 * (size_t)cell + q;
      }
      atomicAdd(addr, v);
    }
  }
};

// Scatter::add_merged with the cell's own terms (SH row, density, merged
// position) added to its LDS slot; the face terms of the neighbouring cells
// outside the chunk go to memory as before.
template <int G, int K>
__device__ __forceinline__ void add_merged_hot(const Scatter<G, K>& sc, HotCells<K>& hc, int at,
                                               const bvh::GradPtrs& gr, V3 gc, float d_sigma, V3 gxyz, bool bprev,
                                               uint32_t prev, V3 gprev, bool bnext, uint32_t next, V3 gnext) {
  constexpr int N = Scatter<G, K>::N;
  const uint32_t lane = threadIdx.x & (G - 1);
  float* row = hc.val[at];
#pragma unroll
  for (int r = 0; r < Scatter<G, K>::R; ++r) {
    const uint32_t item = lane + r * G;
    if (item < 3u * K) {
      atomicAdd(row + item, pick3(gc, item % 3u) * sc.bval[r]);
    } else if (item < (uint32_t)N) {
      const uint32_t q = item - 3u * K;
      const uint32_t part = q == 0 ? 0u : (q < 4 ? 1u : (q < 7 ? 2u : 3u));
      const uint32_t comp = q == 0 ? 0u : (q - 1u) % 3u;
      if (part == 0) {
        atomicAdd(row + 3 * K + 3, d_sigma);
      } else if (part == 1) {
        atomicAdd(row + 3 * K + comp, pick3(gxyz, comp));
      } else if ((part == 2 && bprev) || (part == 3 && bnext)) {
        const uint32_t cell = part == 2 ? prev : next;
        float* base = gr.packed ? reinterpret_cast<float*>(gr.packed + cell) : gr.d_xyz + 3 * (size_t)cell;
        atomicAdd(base + comp, pick3(part == 2 ? gprev : gnext, comp));
      }
    }
  }
}

// The packed (x, y, z, density) gradients of a training launch added to the
// caller's tables.  Atomic: launches on other streams may add to the same
// tables at the same time.
__global__ void k_zero_packed(int64_t n, float4* p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}
__global__ void k_unpack_grads(int64_t n, const float4* __restrict__ p, float* d_density, float* d_xyz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = p[i];
  if (v.w != 0.f) atomicAdd(d_density + i, v.w);
  if (v.x != 0.f) atomicAdd(d_xyz + 3 * i, v.x);
  if (v.y != 0.f) atomicAdd(d_xyz + 3 * i + 1, v.y);
  if (v.z != 0.f) atomicAdd(d_xyz + 3 * i + 2, v.z);
}

// Gradient factors of a plane crossing (trace.rs:1077-1097, radius terms
// dropped); scaled by one reciprocal, which the 1e-4 gradient tolerance allows.
__device__ __forceinline__ void face_jac(V3 ro, V3 rd, V3 cur, V3 adj, float t, V3* d_cur, V3* d_adj) {
  const V3 nrm = sub(adj, cur);
  const float den = dot(rd, nrm);
  const bool ok = isfinite(den) && fabsf(den) > 1e-20f;
  // a gradient, not a cell choice: the hardware reciprocal (~1 ulp) suffices
  const float inv = ok ? __builtin_amdgcn_rcpf(den) : 0.0f;
  const V3 a = add(sub(ro, cur), scl(t, rd));
  const V3 b = sub(sub(adj, ro), scl(t, rd));
  *d_cur = scl(inv, a);
  *d_adj = scl(inv, b);
}

// ---------------------------------------------------------------- forward
// The forward walk of one ray by its group (trace.rs:1680-1820 semantics).
// The first `cap` steps of one ray's forward walk, kept for its replay as one
// 64-byte record per step: (next cell xyz, exit depth t1), (unclamped rgb,
// density), (entry depth t0, transmittance, accumulated r, g before the
// step), (accumulated b, next cell id bits, BV_NO_CELL where no exit face was
// found).  With the state before each step the replay's steps are independent
// of each other (replay_cached).  Step k is written by lane k mod G of the
// ray's group (its values are replicated on every lane).
struct PathCache {
  float4* rec;    // [cap][4]
  uint32_t cap;
};

// Work order of a launch over its n rays: the i-th ray taken is ray
// (i * mult) mod n (bvh::scatter_mult), so the rays that run side by side,
// in one wave or across the device, come from all over the call.
struct WorkOrder {
  uint32_t n;
  uint32_t mult;
};
__device__ __forceinline__ int64_t scattered(const WorkOrder& wo, int64_t i) {
  return (int64_t)(((uint64_t)i * wo.mult) % wo.n);
}

// WIN: integrate only from t_lo on (integration_begin = max(support start,
// integration_start), radfoam_trace.wgsl:159-160); the walk itself still
// starts at t = 0 in the start cell.  Object scenes use it with the object's
// bounding-sphere interval.
template <int G, int U, int EVAL, int DEG, bool BOUNDED, bool CACHE = false, bool WIN = false>
__device__ __forceinline__ void walk_forward(const SceneView& s, const Settings& st, V3 ro, V3 dir,
                                             const Colour<G, DEG>& col, uint32_t start, bool count_bytes,
                                             WalkAcc* a, uint32_t* steps_out, uint32_t* last_out, float* t_end_out,
                                             const PathCache* pc = nullptr, uint32_t* cached = nullptr,
                                             float t_lo = 0.0f, uint32_t* path = nullptr, uint32_t path_len = 0) {
  using Col = Colour<G, DEG>;
  constexpr int K = Col::K;
  constexpr int KM = Col::KM;
  constexpr bool SH = EVAL == BV_EVAL_SH;
  Cell c = start_cell(s, start);
  Prefetch<U, KM> f;
  prefetch<G, U, K, KM, SH>(s, c, &f);
  float t0 = 0.0f;
  uint32_t steps = 0;
  while (steps < st.max_steps) {
    steps += 1;
    // the visited-cell sequence, when asked for (verification)
    if (path != nullptr && steps - 1 < path_len && glane<G>() == 0) path[steps - 1] = c.id;
    if (a->trans <= st.weight_threshold || t0 > st.depth) break;
    float t1;
    Cell nx = c;
    const bool found = coop_exit<G, U, BOUNDED, !CACHE>(s, c, f, ro, dir, t0, st.depth, &t1, &nx);
    // The next cell's loads go out now and land while this cell is integrated.
    Prefetch<U, KM> fn;
    prefetch<G, U, K, KM, SH>(s, found ? nx : c, &fn);
    const float4 cd = SH ? col.eval(f.col) : make_float4(0.f, 0.f, 0.f, f.col[0].w);
    if (pc != nullptr) {
      if (steps - 1 < pc->cap && glane<G>() == ((steps - 1) & (G - 1))) {
        float4* r = pc->rec + 4 * (size_t)(steps - 1);
        r[0] = make_float4(nx.p.x, nx.p.y, nx.p.z, t1);
        r[1] = cd;
        r[2] = make_float4(t0, a->trans, a->r, a->g);
        r[3] = make_float4(a->b, __uint_as_float(found ? nx.id : BV_NO_CELL), 0.f, 0.f);
      }
      *cached = steps;  // steps walked; all of them are recorded while this is <= cap
    }
    if (count_bytes) a->bytes += step_bytes(c.end - c.begin, BOUNDED);
    float seg0, seg1;
    bool seg = support_interval(BOUNDED, ro, dir, c.p, c.r, t0, t1, &seg0, &seg1);
    if (WIN && seg) {
      seg0 = seg0 > t_lo ? seg0 : t_lo;
      seg = seg1 > seg0;
    }
    if (seg && cd.w > 1e-6f) {
      if (SH && count_bytes) a->bytes += 12u * K;
      integrate<EVAL>(st, cd.w, seg0, seg1, c.id, clamp0(mk(cd.x, cd.y, cd.z)), a);
    }
    if (!found) {
      t0 = t1;
      break;
    }
    t0 = t0 > t1 ? t0 : t1;
    c = nx;
    f = fn;
  }
  *steps_out = steps;
  *last_out = c.id;
  *t_end_out = t0;
}

template <int G, int U, int EVAL, int DEG, bool BOUNDED>
__global__ void __launch_bounds__(256) k_trace_coop(SceneView s, RaySource rs, Settings st, Outputs out) {
  const int64_t slot = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (slot >= ray_slots(rs)) return;
  V3 ro, dir;
  uint32_t start;
  int64_t oi;
  if (!fetch_ray(rs, slot, &ro, &dir, &start, &oi)) return;
  WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
  const bool writer = glane<G>() == 0;
  if (!normalise(&dir)) {
    if (writer) store_out(out, oi, a, 0u, start, 0.0f);
    return;
  }
  Colour<G, DEG> col;
  if (EVAL == BV_EVAL_SH) col.init(dir);
  uint32_t steps, last;
  float t_end;
  walk_forward<G, U, EVAL, DEG, BOUNDED>(s, st, ro, dir, col, start, out.stats != nullptr, &a, &steps, &last, &t_end,
                                         nullptr, nullptr, 0.0f,
                                         out.path ? out.path + oi * (int64_t)out.path_len : nullptr, out.path_len);
  if (writer) store_out(out, oi, a, steps, last, t_end);
}

// Persistent k_trace_coop for calls with more rays than the device holds
// groups.  A group whose walk has ended takes the next ray at once instead of
// waiting for the longest walk of its wave: each wave claims a lot of rays
// with one counter atomic (`next_lot`, zero before the launch), and the groups
// without a ray take the lot's next rays in group order (a ballot of the
// groups asking and each one's rank in it).  Every step is walk_forward's, so
// each ray's outputs are the same bit for bit.  The wave leaves once its lot
// and the call's rays are used up and none of its groups walks.
template <int G, int U, int EVAL, int DEG, bool BOUNDED>
__global__ void __launch_bounds__(256) k_trace_pool(SceneView s, RaySource rs, Settings st, Outputs out,
                                                    uint32_t* next_lot) {
  using Col = Colour<G, DEG>;
  constexpr int K = Col::K;
  constexpr int KM = Col::KM;
  constexpr bool SH = EVAL == BV_EVAL_SH;
  constexpr uint32_t kLot = 4u * (64u / G);  // rays per wave claim
  const uint32_t lane = glane<G>();
  const bool leader = lane == 0;
  const uint32_t wl = threadIdx.x & 63u;
  const uint64_t slots = (uint64_t)ray_slots(rs);
  const bool count_bytes = out.stats != nullptr;
  // the wave's lot: slots [pool, pool_end) not taken yet (wave-uniform)
  uint64_t pool = 0, pool_end = 0;
  bool exhausted = false;
  // the group's walk
  bool live = false;
  V3 ro = mk(0.f, 0.f, 0.f), dir = ro;
  int64_t oi = 0;
  Col col;
  Cell c{0u, 0u, 0u, ro, 0.f};
  Prefetch<U, KM> f;
  float t0 = 0.0f;
  uint32_t steps = 0;
  WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
  for (;;) {
    const uint64_t asking = __ballot(leader && !live);
    if (asking != 0ull) {
      if (pool == pool_end && !exhausted) {
        uint32_t b = 0;
        if (wl == 0) b = atomicAdd(next_lot, kLot);
        b = __builtin_amdgcn_readfirstlane(b);
        if ((uint64_t)b >= slots) {
          exhausted = true;
        } else {
          pool = b;
          pool_end = (uint64_t)b + kLot < slots ? (uint64_t)b + kLot : slots;
        }
      }
      const uint32_t n_ask = (uint32_t)__popcll(asking);
      const uint64_t avail = pool_end - pool;
      const uint32_t rank = (uint32_t)__popcll(asking & ((1ull << (wl & ~(uint32_t)(G - 1))) - 1ull));
      const bool take = !live && (uint64_t)rank < avail;
      const uint64_t slot = pool + rank;
      pool += (uint64_t)n_ask < avail ? (uint64_t)n_ask : avail;
      if (take) {
        uint32_t start;
        if (fetch_ray(rs, (int64_t)slot, &ro, &dir, &start, &oi)) {
          a = WalkAcc{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
          if (!normalise(&dir)) {
            if (leader) store_out(out, oi, a, 0u, start, 0.0f);
          } else {
            if (SH) col.init(dir);
            c = start_cell(s, start);
            prefetch<G, U, K, KM, SH>(s, c, &f);
            t0 = 0.0f;
            steps = 0;
            live = true;
          }
        }
      }
    }
    if (__ballot(live) == 0ull) {
      if (exhausted && pool == pool_end) break;
    } else if (live) {
      // one iteration of walk_forward's loop
      bool end = false;
      if (steps >= st.max_steps) {
        end = true;
      } else {
        steps += 1;
        if (out.path != nullptr && steps - 1 < out.path_len && leader)
          out.path[oi * (int64_t)out.path_len + steps - 1] = c.id;
        if (a.trans <= st.weight_threshold || t0 > st.depth) {
          end = true;
        } else {
          float t1;
          Cell nx = c;
          const bool found = coop_exit<G, U, BOUNDED, true>(s, c, f, ro, dir, t0, st.depth, &t1, &nx);
          Prefetch<U, KM> fn;
          prefetch<G, U, K, KM, SH>(s, found ? nx : c, &fn);
          const float4 cd = SH ? col.eval(f.col) : make_float4(0.f, 0.f, 0.f, f.col[0].w);
          if (count_bytes) a.bytes += step_bytes(c.end - c.begin, BOUNDED);
          float seg0, seg1;
          const bool seg = support_interval(BOUNDED, ro, dir, c.p, c.r, t0, t1, &seg0, &seg1);
          if (seg && cd.w > 1e-6f) {
            if (SH && count_bytes) a.bytes += 12u * K;
            integrate<EVAL>(st, cd.w, seg0, seg1, c.id, clamp0(mk(cd.x, cd.y, cd.z)), &a);
          }
          if (!found) {
            t0 = t1;
            end = true;
          } else {
            t0 = t0 > t1 ? t0 : t1;
            c = nx;
            f = fn;
          }
        }
      }
      if (end) {
        if (leader) store_out(out, oi, a, steps, c.id, t0);
        live = false;
      }
    }
  }
}

// ---------------------------------------------------------------- recorder
template <int G, int U, bool BOUNDED, class RS>
__global__ void __launch_bounds__(256) k_record_coop(SceneView s, RS rs, uint32_t L, float depth, float max_dt,
                                                     uint32_t* cells, uint32_t* next_cells, float* dts, float* mask,
                                                     uint32_t* status) {
  const int64_t slot = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (slot >= ray_slots(rs)) return;
  V3 ro, dir;
  uint32_t start;
  int64_t oi;
  const bool writer = glane<G>() == 0;
  if (!fetch_ray(rs, slot, &ro, &dir, &start, &oi, writer)) return;
  if (!normalise(&dir)) {
    if (writer) status[oi] = 0u;
    return;
  }
  const int64_t rowp = oi * (int64_t)L;
  Cell c = start_cell(s, start);
  Prefetch<U, 1> f;
  prefetch<G, U, 1, 1, false>(s, c, &f);
  float t0 = 0.0f;
  uint32_t count = 0;
  bool truncated = false;
  // Lane j of the group holds step base + j of the current chunk of G steps;
  // a full chunk goes out as G contiguous values per stream (one store
  // instruction per stream for G steps, instead of four one-lane stores a
  // step, each to its own line: those took half the recorder's time).
  const uint32_t lane = glane<G>();
  uint32_t h_cell = 0u, h_next = 0u;
  float h_dt = 0.0f;
  auto store_chunk = [&](uint32_t base, uint32_t n) {
    if (lane < n) {
      const int64_t at = rowp + base + lane;
      cells[at] = h_cell;
      next_cells[at] = h_next;
      dts[at] = h_dt;
      if (mask) mask[at] = 1.0f;
    }
  };
  for (uint32_t step = 0; step < L; ++step) {
    float t1;
    Cell nx = c;
    const bool found = coop_exit<G, U, BOUNDED, false>(s, c, f, ro, dir, t0, depth, &t1, &nx);
    Prefetch<U, 1> fn;
    prefetch<G, U, 1, 1, false>(s, found ? nx : c, &fn);
    if (t1 <= t0) break;
    const uint32_t nxt = found ? nx.id : c.id;
    float dt = t1 - t0;
    dt = dt < max_dt ? dt : max_dt;
    if (lane == (count & (G - 1))) {
      h_cell = c.id;
      h_next = nxt;
      h_dt = dt;
    }
    count += 1;
    if ((count & (G - 1)) == 0) store_chunk(count - G, G);
    if (!found) break;
    if (step + 1 == L) truncated = true;
    t0 = t1;
    c = nx;
    f = fn;
  }
  const uint32_t tail = count & (G - 1);
  if (tail) store_chunk(count - tail, tail);
  if (writer) status[oi] = count | (truncated ? BV_PATH_TRUNCATED : 0u);
}

// ---------------------------------------------------------------- adjoint

// Replay of the forward walk (same arithmetic, same cells) applying, per
// integrated segment, the closed-form gradients listed in bv_adjoint.hip.  C is
// the forward's rgba, g = dL/d(rgba).  The scatter is striped over the group
// (Scatter above); each step adds its cell's terms and its face terms for the
// previous and next cells at once.  The fused kernels use it for walks longer
// than their path cache.
template <int G, int U, int DEG>
__device__ __forceinline__ void walk_adjoint(const SceneView& s, const Settings& st, V3 ro, V3 dir,
                                             const Colour<G, DEG>& col, uint32_t start, float4 C, float4 g,
                                             const bvh::GradPtrs& gr) {
  constexpr int K = Colour<G, DEG>::K;
  constexpr int KM = Colour<G, DEG>::KM;
  const float Tfin = 1.0f - C.w;
  Scatter<G, K> sc;
  sc.init(dir);
  float trans = 1.0f, pr = 0.f, pg = 0.f, pb = 0.f;
  Cell c = start_cell(s, start);
  Prefetch<U, KM> f;
  prefetch<G, U, K, KM, true>(s, c, &f);
  V3 pp = c.p;
  uint32_t prev = c.id;
  float t0 = 0.0f;
  uint32_t steps = 0;
  bool have_exit = false;
  V3 exit_dc = mk(0.f, 0.f, 0.f), exit_dn = exit_dc;
  float exit_t = 0.0f;
  while (steps < st.max_steps) {
    steps += 1;
    if (trans <= st.weight_threshold || t0 > st.depth) break;
    float t1;
    Cell nx = c;
    bool found;
    float4 cd;
    Prefetch<U, KM> fn;
    found = coop_exit<G, U, false, false>(s, c, f, ro, dir, t0, st.depth, &t1, &nx);
    prefetch<G, U, K, KM, true>(s, found ? nx : c, &fn);
    cd = col.eval(f.col);
    const float sd = cd.w;
    const bool integ = t1 > t0 && sd > 1e-6f;
    const float dt = t1 - t0;
    const float alpha = 1.0f - expf(-sd * dt);
    const float w = trans * alpha;
    const float cr = cd.x > 0.f ? cd.x : 0.f, cg = cd.y > 0.f ? cd.y : 0.f, cb = cd.z > 0.f ? cd.z : 0.f;
    const float npr = pr + w * cr, npg = pg + w * cg, npb = pb + w * cb;
    const float tn = trans * (1.0f - alpha);
    const float gtau = g.x * (tn * cr - (C.x - npr)) + g.y * (tn * cg - (C.y - npg)) +
                       g.z * (tn * cb - (C.z - npb)) + g.w * Tfin;
    const float d_sigma = integ ? gtau * dt : 0.f;
    const float d_dt = integ ? gtau * sd : 0.f;
    const V3 gc = mk(integ && cd.x > 0.f ? g.x * w : 0.f, integ && cd.y > 0.f ? g.y * w : 0.f,
                     integ && cd.z > 0.f ? g.z * w : 0.f);
    // Positions through dt = t1 - t0 (exit face minus entry face).  The entry
    // face is the previous step's exit face at the same depth, whose Jacobian
    // that step already formed.
    V3 gcur = mk(0.f, 0.f, 0.f), gnext = gcur, gprev = gcur;
    V3 xc = gcur, xn = gcur;
    if (found) {
      face_jac(ro, dir, c.p, nx.p, t1, &xc, &xn);
      gcur = scl(d_dt, xc);
      gnext = scl(d_dt, xn);
    }
    if (prev != c.id) {
      V3 dp, dc;
      if (have_exit && t0 == exit_t) {
        dp = exit_dc;
        dc = exit_dn;
      } else {
        face_jac(ro, dir, pp, c.p, t0, &dp, &dc);
      }
      gcur = sub(gcur, scl(d_dt, dc));
      gprev = scl(-d_dt, dp);
    }
    have_exit = found;
    exit_dc = xc;
    exit_dn = xn;
    exit_t = t1;
    sc.add_merged(gr, c.id, gc, d_sigma, gcur, prev != c.id, prev, gprev, found, nx.id, gnext);
    if (integ) {
      pr = npr;
      pg = npg;
      pb = npb;
      trans = tn;
    }
    if (!found) break;
    t0 = t0 > t1 ? t0 : t1;
    prev = c.id;
    pp = c.p;
    c = nx;
    f = fn;
  }
}

// Replay of a walk whose every step is in the path cache.  With the forward's
// state before each step (entry depth, transmittance, accumulated colour) the
// steps do not depend on each other, so lane j of the group takes step
// base + j: it forms the step's blend terms and both face Jacobians (the same
// operations on the same FP32 values as the sequential replay, walk_adjoint)
// and stages the step's gradient values in LDS; the group then adds the staged
// steps in order with the coalesced scatter.  `stage` is the group's [G][4]
// rows of LDS.
// Steps [base, base + G) of a recorded walk of n_steps steps, step-parallel:
// lane j takes step base + j (rec_at(k) is step k's 64 B record).
template <int G, int DEG, class RecAt>
__device__ __forceinline__ void replay_chunk(V3 ro, V3 dir, uint32_t start, float4 p0, float4 C, float Tfin,
                                             float4 g, const bvh::GradPtrs& gr,
                                             Scatter<G, Colour<G, DEG>::K>& sc, RecAt&& rec_at, uint32_t base,
                                             uint32_t n_steps, float4 (*stage)[4],
                                             HotCells<Colour<G, DEG>::K>* hot = nullptr) {
  const uint32_t lane = glane<G>();
  const uint32_t k = base + lane;
  const bool mine = k < n_steps;
  const uint32_t kk = mine ? k : n_steps - 1;
  const float4* r = rec_at(kk);
  const float4 R0 = r[0], R1 = r[1], R2 = r[2], R3 = r[3];
  // the step's cell is the previous step's next cell, its previous cell the
  // one before (the start cell for the first steps)
  float4 Q0 = p0, P0 = p0;
  uint32_t cid = start, pid = start;
  if (kk >= 1) {
    const float4* q = rec_at(kk - 1);
    Q0 = q[0];
    cid = __float_as_uint(q[3].y);
  }
  if (kk >= 2) {
    const float4* q = rec_at(kk - 2);
    P0 = q[0];
    pid = __float_as_uint(q[3].y);
  }
  const V3 cp = mk(Q0.x, Q0.y, Q0.z), pp = mk(P0.x, P0.y, P0.z), np = mk(R0.x, R0.y, R0.z);
  const uint32_t nid = __float_as_uint(R3.y);
  const bool found = mine && nid != BV_NO_CELL;
  const float t1 = R0.w, t0 = R2.x, trans = R2.y;
  const float4 cd = R1;
  const float sd = cd.w;
  const bool integ = mine && t1 > t0 && sd > 1e-6f;
  const float dt = t1 - t0;
  const float alpha = 1.0f - expf(-sd * dt);
  const float w = trans * alpha;
  const float cr = cd.x > 0.f ? cd.x : 0.f, cg = cd.y > 0.f ? cd.y : 0.f, cb = cd.z > 0.f ? cd.z : 0.f;
  const float npr = R2.z + w * cr, npg = R2.w + w * cg, npb = R3.x + w * cb;
  const float tn = trans * (1.0f - alpha);
  const float gtau = g.x * (tn * cr - (C.x - npr)) + g.y * (tn * cg - (C.y - npg)) +
                     g.z * (tn * cb - (C.z - npb)) + g.w * Tfin;
  const float d_sigma = integ ? gtau * dt : 0.f;
  const float d_dt = integ ? gtau * sd : 0.f;
  const V3 gc = mk(integ && cd.x > 0.f ? g.x * w : 0.f, integ && cd.y > 0.f ? g.y * w : 0.f,
                   integ && cd.z > 0.f ? g.z * w : 0.f);
  // positions through dt = t1 - t0: the exit face, then the entry face
  V3 gcur = mk(0.f, 0.f, 0.f), gnext = gcur, gprev = gcur;
  if (found) {
    V3 xc, xn;
    face_jac(ro, dir, cp, np, t1, &xc, &xn);
    gcur = scl(d_dt, xc);
    gnext = scl(d_dt, xn);
  }
  if (mine && pid != cid) {
    V3 dp, dc;
    face_jac(ro, dir, pp, cp, t0, &dp, &dc);
    gcur = sub(gcur, scl(d_dt, dc));
    gprev = scl(-d_dt, dp);
  }
  const uint32_t cell = mine ? cid : start;
  stage[lane][0] = make_float4(gc.x, gc.y, gc.z, d_sigma);
  stage[lane][1] = make_float4(gcur.x, gcur.y, gcur.z, __uint_as_float(cell));
  stage[lane][2] = make_float4(gnext.x, gnext.y, gnext.z, __uint_as_float(found ? nid : cell));
  stage[lane][3] = make_float4(gprev.x, gprev.y, gprev.z, __uint_as_float(mine ? pid : start));
  // the rows are read by the other lanes of the group (one wave)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t cnt = n_steps - base < (uint32_t)G ? n_steps - base : (uint32_t)G;
  // Step j's cell is step j - 1's next cell and step j + 1's previous one,
  // so their face terms for it (gnext of j - 1, gprev of j + 1) join its own
  // before the atomics; the chunk's ends add the terms that reach past it.
  for (uint32_t j = 0; j < cnt; ++j) {
    const float4 A = stage[j][0], B = stage[j][1], D = stage[j][2], E = stage[j][3];
    const uint32_t cl = __float_as_uint(B.w), nx = __float_as_uint(D.w), pv = __float_as_uint(E.w);
    V3 gx = mk(B.x, B.y, B.z);
    if (j > 0) {
      const float4 Dp = stage[j - 1][2];
      gx = add(gx, mk(Dp.x, Dp.y, Dp.z));
    }
    if (j + 1 < cnt) {
      const float4 En = stage[j + 1][3];
      gx = add(gx, mk(En.x, En.y, En.z));
    }
    if (hot != nullptr && base + j < kHotSteps) {
      int at = -1;
      if (lane == 0) at = hot->find(cl);
      at = __shfl(at, 0, G);
      if (at >= 0) {
        add_merged_hot<G>(sc, *hot, at, gr, mk(A.x, A.y, A.z), A.w, gx, j == 0 && pv != cl, pv,
                          mk(E.x, E.y, E.z), j + 1 == cnt && nx != cl, nx, mk(D.x, D.y, D.z));
        continue;
      }
    }
    sc.add_merged(gr, cl, mk(A.x, A.y, A.z), A.w, gx, j == 0 && pv != cl, pv, mk(E.x, E.y, E.z),
                  j + 1 == cnt && nx != cl, nx, mk(D.x, D.y, D.z));
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G, int DEG>
__device__ __forceinline__ void replay_cached(const SceneView& s, const Settings& st, V3 ro, V3 dir, uint32_t start,
                                              float4 C, float4 g, const bvh::GradPtrs& gr, const PathCache& pc,
                                              uint32_t n_steps, float4 (*stage)[4],
                                              HotCells<Colour<G, DEG>::K>* hot = nullptr) {
  constexpr int K = Colour<G, DEG>::K;
  const float Tfin = 1.0f - C.w;
  Scatter<G, K> sc;
  sc.init(dir);
  const CellRec c0 = s.cells[start];
  const float4 p0 = make_float4(c0.x, c0.y, c0.z, 0.f);
  for (uint32_t base = 0; base < n_steps; base += G)
    replay_chunk<G, DEG>(ro, dir, start, p0, C, Tfin, g, gr, sc,
                         [&](uint32_t k) { return pc.rec + 4 * (size_t)k; }, base, n_steps, stage, hot);
}

__global__ void k_reset_counter(uint32_t* counter) { *counter = 0u; }

// What a fused launch takes and gives per ray: the L2 loss (target colours
// in; colours and the summed loss out, each optional) or the adjoint of a
// caller's forward colours and loss gradients ([count][4] each).
struct RayTerms {
  const float* target;
  float* rgba_out;
  float* loss;
  const float* rgba_fwd;
  const float* d_rgba;
};

// Persistent: the grid holds at most as many blocks as fit on the device at
// once, and each wave claims its next 64 / G rays from `next_ray` (reset to
// 0 before the launch) until the rays run out, so one launch keeps every SIMD
// busy however many batches it carries and however unequal their walks.  Rays
// are taken in a scattered order: rays a caller grouped by camera cross the
// same cells, and side by side (above all within one wave) their gradient
// atomics collide.  The claim is wave-uniform, so every wave leaves the loop
// at the same claim.
// Each group owns one region of the path cache for all the rays it walks.
// ADJ: the adjoint of given forward colours (rgba_fwd) and loss gradients
// (d_rgba) instead of the L2 loss; the walk only fills the path cache.
template <int G, int U, int DEG, bool TRACED = false, bool ADJ = false>
__global__ void __launch_bounds__(256) k_train_l2_coop(SceneView s, RaySource rs, Settings st, RayTerms t,
                                                       bvh::GradPtrs gr, PathCache paths, uint32_t* next_ray,
                                                       WorkOrder wo, uint32_t* path_out, uint32_t path_len) {
  // replay staging: one [G][4] float4 row block per group of the 256 lanes
  __shared__ float4 stage_rows[256][4];
  __shared__ HotCells<Colour<G, DEG>::K> hot;
  constexpr uint32_t kRaysPerWave = 64 / G;
  hot.clear();
  __syncthreads();
  // this group's rows of the path cache
  const int64_t group = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const PathCache pc{paths.rec + 4 * group * paths.cap, paths.cap};
  float l = 0.f;
  for (;;) {
    uint32_t claim = 0;
    if ((threadIdx.x & 63) == 0) claim = atomicAdd(next_ray, 1u);
    claim = __builtin_amdgcn_readfirstlane(claim);
    const int64_t i = (int64_t)claim * kRaysPerWave + (threadIdx.x & 63) / G;
    if ((int64_t)claim * kRaysPerWave >= rs.count) break;
    // the group's ray, in the scattered order
    // (a guard, not `continue`: one back-edge keeps the walks' loops unsplit)
    const int64_t slot = i < rs.count ? scattered(wo, i) : rs.count;
    uint32_t cached = 0;
    V3 ro, dir;
    uint32_t start;
    int64_t oi;
    if (slot < rs.count && fetch_ray(rs, slot, &ro, &dir, &start, &oi)) {
      WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
      const bool ok = normalise(&dir);
      Colour<G, DEG> col;
      if (ok) {
        col.init(dir);
        uint32_t steps, last;
        float t_end;
        walk_forward<G, U, BV_EVAL_SH, DEG, false, true>(
            s, st, ro, dir, col, start, false, &a, &steps, &last, &t_end, &pc, &cached, 0.0f,
            TRACED ? path_out + oi * (int64_t)path_len : nullptr, TRACED ? path_len : 0u);
      }
      float4 C, gl;
      if (ADJ) {
        C = reinterpret_cast<const float4*>(t.rgba_fwd)[oi];
        gl = reinterpret_cast<const float4*>(t.d_rgba)[oi];
      } else {
        C = make_float4(a.r, a.g, a.b, 1.0f - a.trans);
        const float er = C.x - t.target[3 * oi], eg = C.y - t.target[3 * oi + 1], eb = C.z - t.target[3 * oi + 2];
        if (glane<G>() == 0) {
          if (t.rgba_out) reinterpret_cast<float4*>(t.rgba_out)[oi] = C;
          l += (er * er + eg * eg) + eb * eb;
        }
        gl = make_float4(2.f * er, 2.f * eg, 2.f * eb, 0.f);
      }
      // the group's lanes wrote the cache rows; make them visible to the group
      // (one wave) before it reads them back
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (ok && cached <= pc.cap)
        replay_cached<G, DEG>(s, st, ro, dir, start, C, gl, gr, pc, cached,
                              &stage_rows[(threadIdx.x / G) * G], &hot);
      else if (ok)
        walk_adjoint<G, U, DEG>(s, st, ro, dir, col, start, C, gl, gr);
    }
  }
  // every wave has left the loop (its exit is wave-uniform)
  __syncthreads();
  hot.flush(gr);
  if (!ADJ && t.loss) {
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(t.loss, l);
  }
}

constexpr int unroll_for(int G) { return G >= 32 ? 1 : (G >= 8 ? 2 : (G >= 4 ? 4 : 8)); }

// ---------------------------------------------------------------- relighting
template <int U>
struct PrefetchL {
  float4 a[U], b[U];
  float4 den;  // colour row component 0 (density in .w)
  float4 mat;  // albedo rgb + roughness
};

template <int G, int U, int K>
__device__ __forceinline__ void prefetch_l(const SceneView& s, const Cell& c, PrefetchL<U>* f) {
  const uint32_t lane = glane<G>();
  const float4* er = reinterpret_cast<const float4*>(s.erec);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t j = c.begin + lane + u * G;
    const uint32_t jj = j < c.end ? j : c.begin;
    f->a[u] = er[2 * (size_t)jj];
    f->b[u] = er[2 * (size_t)jj + 1];
  }
  f->den = reinterpret_cast<const float4*>(s.shd)[(size_t)c.id * K];
  f->mat = s.material[c.id];
}

// Threads of the relighting kernels (k_relight_coop, k_serve): each keeps its
// 3 * BV_MAX_LIGHTS per-light sums in its own LDS column.
constexpr int kRelightStride = 256;

// Relit walk of one ray by its group: the forward walk's cells and weights,
// per-light sums of shaded segments, then one shadow walk per light from the
// depth point (start cell = the peak cell) when visibility is on.
template <int G, int U, int DEG, bool BOUNDED>
__device__ __forceinline__ float4 walk_relight(const SceneView& s, const Settings& st, const bv_lighting& lit, V3 ro,
                                               V3 dir, uint32_t start, float* sums) {
  constexpr int K = (DEG + 1) * (DEG + 1);
  const int nl = lit.count < BV_MAX_LIGHTS ? (int)lit.count : BV_MAX_LIGHTS;
  // the per-light sums live in the thread's LDS column (`sums`, stride
  // kRelightStride): indexed by the run-time light count they would sit in
  // scratch
#pragma unroll
  for (int q = 0; q < 3 * BV_MAX_LIGHTS; ++q) sums[q * kRelightStride] = 0.f;
  const V3 view = mk(-dir.x, -dir.y, -dir.z);
  WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
  Cell c = start_cell(s, start);
  PrefetchL<U> f;
  prefetch_l<G, U, K>(s, c, &f);
  float t0 = 0.0f;
  uint32_t steps = 0;
  while (steps < st.max_steps) {
    steps += 1;
    if (a.trans <= st.weight_threshold || t0 > st.depth) break;
    float t1;
    Cell nx = c;
    Prefetch<U, 1> edges;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      edges.a[u] = f.a[u];
      edges.b[u] = f.b[u];
    }
    const bool found = coop_exit<G, U, BOUNDED>(s, c, edges, ro, dir, t0, st.depth, &t1, &nx);
    PrefetchL<U> fn;
    prefetch_l<G, U, K>(s, found ? nx : c, &fn);
    const float sd = f.den.w;
    float seg0, seg1;
    if (support_interval(BOUNDED, ro, dir, c.p, c.r, t0, t1, &seg0, &seg1) && sd > 1e-6f) {
      const float dt = seg1 - seg0;
      const float alpha = 1.0f - expf(-sd * dt);
      const float w = a.trans * alpha;
      const V3 x = add(ro, scl(seg0, dir));
      const V3 nrm = sub(x, c.p);
      for (int l = 0; l < nl; ++l) {
        const V3 col = shade_point_light(lit.lights[l], x, nrm, view, f.mat, lit.specular_f0);
        sums[(3 * l) * kRelightStride] += w * col.x;
        sums[(3 * l + 1) * kRelightStride] += w * col.y;
        sums[(3 * l + 2) * kRelightStride] += w * col.z;
      }
      if (w > a.peak) {
        a.peak = w;
        a.depth_mode = 0.5f * (seg0 + seg1);
        a.peak_point = c.id;
      }
      a.trans *= 1.0f - alpha;
    }
    if (!found) break;
    t0 = t0 > t1 ? t0 : t1;
    c = nx;
    f = fn;
  }
  float4 rgba = make_float4(0.f, 0.f, 0.f, 1.0f - a.trans);
  const bool shadows = a.peak_point != BV_NO_CELL;
  const V3 p = add(ro, scl(a.depth_mode, dir));
  for (int l = 0; l < nl; ++l) {
    float vis = 1.0f;
    // a visibility ray for every light (lit.visibility) or for the lights that ask for one
    if (shadows && (lit.visibility != 0 || lit.lights[l].visibility != 0)) {
      V3 to = sub(mk(lit.lights[l].position[0], lit.lights[l].position[1], lit.lights[l].position[2]), p);
      const float dist = sqrtf(dot(to, to));
      if (dist > 0.0f && normalise(&to)) {
        Settings ss = st;
        ss.depth = dist;
        ss.eval_mode = BV_EVAL_OPACITY;
        WalkAcc sh{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
        Colour<G, DEG> none;
        uint32_t stp, last;
        float te;
        walk_forward<G, U, BV_EVAL_OPACITY, DEG, BOUNDED>(s, ss, p, to, none, a.peak_point, false, &sh, &stp, &last,
                                                          &te);
        vis = sh.trans;
      }
    }
    rgba.x += vis * sums[(3 * l) * kRelightStride];
    rgba.y += vis * sums[(3 * l + 1) * kRelightStride];
    rgba.z += vis * sums[(3 * l + 2) * kRelightStride];
  }
  return rgba;
}

template <int G, int U, int DEG, bool BOUNDED>
__global__ void __launch_bounds__(kRelightStride) k_relight_coop(SceneView s, RaySource rs, Settings st,
                                                                 bv_lighting lit, float* rgba_out) {
  __shared__ float sums[3 * BV_MAX_LIGHTS * kRelightStride];
  const int64_t slot = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (slot >= ray_slots(rs)) return;
  V3 ro, dir;
  uint32_t start;
  int64_t oi;
  if (!fetch_ray(rs, slot, &ro, &dir, &start, &oi)) return;
  float4 rgba = make_float4(0.f, 0.f, 0.f, 0.f);
  if (normalise(&dir)) rgba = walk_relight<G, U, DEG, BOUNDED>(s, st, lit, ro, dir, start, &sums[threadIdx.x]);
  if (glane<G>() == 0) reinterpret_cast<float4*>(rgba_out)[oi] = rgba;
}

template <int G>
void launch_relight_g(const SceneView& v, const RaySource& rs, const Settings& st, const bv_lighting& lit,
                      float* rgba_out, hipStream_t stream) {
  constexpr int U = unroll_for(G);
  const int64_t threads = ray_slots(rs) * G;
  const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
#define BV_R(D)                                                                                              \
  if (v.bounded)                                                                                             \
    hipLaunchKernelGGL((k_relight_coop<G, U, D, true>), grid, block, 0, stream, v, rs, st, lit, rgba_out);  \
  else                                                                                                       \
    hipLaunchKernelGGL((k_relight_coop<G, U, D, false>), grid, block, 0, stream, v, rs, st, lit, rgba_out);
  switch (v.sh_degree) {
    case 0: BV_R(0) break;
    case 1: BV_R(1) break;
    case 2: BV_R(2) break;
    default: BV_R(3) break;
  }
#undef BV_R
}


// One trace launch: the persistent pooled kernel when the call has more rays
// than the device holds groups of the kernel (next_lot given), else one
// group per ray.
template <int G, int U, int EV, int D, bool BND>
void launch_trace_k(const SceneView& v, const RaySource& rs, const Settings& st, const Outputs& out,
                    uint32_t* next_lot, int cus, hipStream_t stream) {
  const int64_t need = (ray_slots(rs) * G + 255) / 256;
  if (next_lot != nullptr) {
    static const int per_cu = [] {
      int b = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_trace_pool<G, U, EV, D, BND>, 256, 0) != hipSuccess ||
          b < 1)
        b = 1;
      return b;
    }();
    const int64_t resident = (int64_t)std::max(cus, 1) * per_cu;
    if (need > resident) {
      hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, next_lot);
      hipLaunchKernelGGL((k_trace_pool<G, U, EV, D, BND>), dim3((unsigned)resident), dim3(256), 0, stream, v, rs,
                         st, out, next_lot);
      return;
    }
  }
  hipLaunchKernelGGL((k_trace_coop<G, U, EV, D, BND>), dim3((unsigned)need), dim3(256), 0, stream, v, rs, st, out);
}

template <int G>
void launch_trace_g(const SceneView& v, const RaySource& rs, const Settings& st, const Outputs& out,
                    uint32_t* next_lot, int cus, hipStream_t stream) {
  constexpr int U = unroll_for(G);
  const bool bnd = v.bounded != 0;
  const int deg = st.eval_mode == BV_EVAL_SH ? v.sh_degree : 0;
#define BV_T(EV, D)                                                         \
  if (bnd)                                                                  \
    launch_trace_k<G, U, EV, D, true>(v, rs, st, out, next_lot, cus, stream); \
  else                                                                      \
    launch_trace_k<G, U, EV, D, false>(v, rs, st, out, next_lot, cus, stream);
  if (st.eval_mode == BV_EVAL_OPACITY) {
    // the density sits at index 3K of the row, so the degree still matters
    switch (v.sh_degree) {
      case 0: BV_T(BV_EVAL_OPACITY, 0) break;
      case 1: BV_T(BV_EVAL_OPACITY, 1) break;
      case 2: BV_T(BV_EVAL_OPACITY, 2) break;
      default: BV_T(BV_EVAL_OPACITY, 3) break;
    }
  } else if (st.eval_mode == BV_EVAL_CONSTANT) {
    switch (v.sh_degree) {
      case 0: BV_T(BV_EVAL_CONSTANT, 0) break;
      case 1: BV_T(BV_EVAL_CONSTANT, 1) break;
      case 2: BV_T(BV_EVAL_CONSTANT, 2) break;
      default: BV_T(BV_EVAL_CONSTANT, 3) break;
    }
  } else {
    switch (deg) {
      case 0: BV_T(BV_EVAL_SH, 0) break;
      case 1: BV_T(BV_EVAL_SH, 1) break;
      case 2: BV_T(BV_EVAL_SH, 2) break;
      default: BV_T(BV_EVAL_SH, 3) break;
    }
  }
#undef BV_T
}

template <int G, class RS>
void launch_record_g(const SceneView& v, const RS& rs, uint32_t L, float depth, float max_dt, uint32_t* cells,
                     uint32_t* next_cells, float* dts, float* mask, uint32_t* status, hipStream_t stream) {
  constexpr int U = unroll_for(G);
  const int64_t threads = ray_slots(rs) * G;
  hipLaunchKernelGGL((k_record_coop<G, U, false, RS>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, stream,
                     v, rs, L, depth, max_dt, cells, next_cells, dts, mask, status);
}

// Blocks of the fused training kernel one CU holds at once (registers and LDS
// permitting), from the occupancy calculator; 1 if it cannot say.
template <int G, int DEG, bool ADJ>
int train_blocks_per_cu() {
  static const int blocks = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_train_l2_coop<G, unroll_for(G), DEG, false, ADJ>, 256,
                                                     0) != hipSuccess ||
        b < 1)
      b = 1;
    return b;
  }();
  return blocks;
}

template <int G>
int train_blocks_per_cu_g(int deg, bool adj) {
  switch (deg) {
    case 0: return adj ? train_blocks_per_cu<G, 0, true>() : train_blocks_per_cu<G, 0, false>();
    case 1: return adj ? train_blocks_per_cu<G, 1, true>() : train_blocks_per_cu<G, 1, false>();
    case 2: return adj ? train_blocks_per_cu<G, 2, true>() : train_blocks_per_cu<G, 2, false>();
    default: return adj ? train_blocks_per_cu<G, 3, true>() : train_blocks_per_cu<G, 3, false>();
  }
}

template <int G>
void launch_train_g(const SceneView& v, const RaySource& rs, const Settings& st, const RayTerms& t, bool adj,
                    const bvh::GradPtrs& gr, const PathCache& pc, uint32_t* next_ray, unsigned blocks,
                    uint32_t* path, uint32_t path_len, hipStream_t stream) {
  constexpr int U = unroll_for(G);
  const dim3 grid(blocks), block(256);
  const WorkOrder wo{(uint32_t)rs.count, bvh::scatter_mult((uint32_t)rs.count)};
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, stream, next_ray);
#define BV_TR(D)                                                                                                   \
  if (adj)                                                                                                         \
    hipLaunchKernelGGL((k_train_l2_coop<G, U, D, false, true>), grid, block, 0, stream, v, rs, st, t, gr, pc,     \
                       next_ray, wo, nullptr, 0u);                                                                 \
  else if (path)                                                                                                   \
    hipLaunchKernelGGL((k_train_l2_coop<G, U, D, true>), grid, block, 0, stream, v, rs, st, t, gr, pc, next_ray,  \
                       wo, path, path_len);                                                                        \
  else                                                                                                             \
    hipLaunchKernelGGL((k_train_l2_coop<G, U, D, false>), grid, block, 0, stream, v, rs, st, t, gr, pc, next_ray, \
                       wo, nullptr, 0u);
  switch (v.sh_degree) {
    case 0: BV_TR(0) break;
    case 1: BV_TR(1) break;
    case 2: BV_TR(2) break;
    default: BV_TR(3) break;
  }
#undef BV_TR
}

// ---------------------------------------------------------------- serving
// Persistent serving kernel over the scheduler's tile ring (ServeQueue): block
// 0 mirrors the host's counters, every other block takes one 8x8 tile per ring
// entry (256 threads = 64 rays x 4 lanes), writes its pixels to the request's
// result buffer in pinned host memory and marks the entry done.

template <int MODE, int DEG, bool BOUNDED>
__global__ void __launch_bounds__(kRelightStride) k_serve(ServeQueue q) {
  constexpr int G = 4;
  constexpr int U = unroll_for(G);
  __shared__ float sums[MODE == BV_EVAL_RELIGHT ? 3 * BV_MAX_LIGHTS * kRelightStride : 1];
  __shared__ ServeReq req;
  __shared__ uint64_t entry_s;
  __shared__ uint32_t code_s;
  if (blockIdx.x == 0) {
    serve_mirror(q);
    return;
  }
  const uint32_t ray = threadIdx.x / G;  // 0..63 within the tile
  for (;;) {
    serve_claim(q, &entry_s, &code_s);
    const uint32_t code = code_s;
    const uint64_t entry = entry_s;
    if (code == kServeNoTile) break;
    serve_load_req(q, code >> 20, &req);
    const SceneView s = uniform_copy(req.s);
    const RaySource rs = uniform_copy(req.rs);
    const Settings st = uniform_copy(req.st);
    float* const out = uniform_copy(req.out);
    V3 ro, dir;
    uint32_t start;
    int64_t oi;
    const int64_t slot = (int64_t)(code & 0xFFFFFu) * 64 + ray;
    if (fetch_ray(rs, slot, &ro, &dir, &start, &oi)) {
      float4 rgba = make_float4(0.f, 0.f, 0.f, 0.f);
      if (normalise(&dir)) {
        if (MODE == BV_EVAL_RELIGHT) {
          rgba = walk_relight<G, U, DEG, BOUNDED>(s, st, req.lit, ro, dir, start, &sums[threadIdx.x]);
        } else {
          WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
          Colour<G, DEG> col;
          col.init(dir);
          uint32_t steps, last;
          float t_end;
          walk_forward<G, U, BV_EVAL_SH, DEG, BOUNDED>(s, st, ro, dir, col, start, false, &a, &steps, &last, &t_end);
          rgba = make_float4(a.r, a.g, a.b, 1.0f - a.trans);
        }
      }
      if (glane<G>() == 0) reinterpret_cast<float4*>(out)[oi] = rgba;
    }
    serve_done(q, entry);
  }
}

// ---------------------------------------------------------------- object scenes
// ray_sphere_intersect (scene_trace_core.wgsl:14-41): t_near clamped to 0, no
// hit when the far root is behind the origin.
__device__ __forceinline__ bool sphere_hit(V3 ro, V3 rd, const float (&c)[3], float r, float* tn, float* tf) {
  const V3 oc = sub(ro, mk(c[0], c[1], c[2]));
  const float a = dot(rd, rd);
  const float b = 2.0f * dot(oc, rd);
  const float cc = dot(oc, oc) - r * r;
  const float disc = b * b - 4.0f * a * cc;
  if (disc < 0.0f) return false;
  const float sq = sqrtf(disc);
  const float t1 = (-b - sq) / (2.0f * a);
  const float t2 = (-b + sq) / (2.0f * a);
  if (t2 < 0.0f) return false;
  *tn = t1 > 0.0f ? t1 : 0.0f;
  *tf = t2;
  return true;
}

__device__ __forceinline__ V3 xform_point(const float (&m)[12], V3 p) {
  return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
            m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
__device__ __forceinline__ V3 xform_dir(const float (&m)[12], V3 d) {
  return mk(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
            m[8] * d.x + m[9] * d.y + m[10] * d.z);
}

// trace_scene (scene_trace_core.wgsl:406-465), G lanes per ray: the bounding
// spheres form a linear software TLAS (at most BV_MAX_OBJECTS entries, held in
// registers), objects are visited in (t_near, index) order, each one walked in
// its own space with the forward walk restricted to [t_near, t_far].  The
// group's lanes hold the same ray state and visit the same objects; they split
// each cell's edge records in the walk, and lane 0 writes.  Held to 4 waves
// per SIMD: at 141 VGPRs it ran 3, and 128 with 36 B spilled per lane traces
// a 1024x768 frame of four 1M-cell instances in 2.0 ms instead of 2.38.
template <int G, int U, int EVAL, int DEG, bool BOUNDED>
__global__ void __launch_bounds__(256, 4) k_trace_objects(ObjectSet os, const float* __restrict__ origins,
                                                       const float* __restrict__ dirs, int64_t count, Settings st,
                                                       float* __restrict__ rgba) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  if (i >= count) return;
  const V3 ro = mk(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]);
  V3 d = mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, T = 1.0f;
  if (normalise(&d)) {
    // find_next_object_hit (:55-74): each visit re-tests the spheres (their
    // data is uniform: scalar loads) and takes the least (t_near, index) after
    // the cursor, so no per-object interval stays live across the walks.
    int cur_k = -1;
    float cur_t = 0.0f;
    while (!(T < st.weight_threshold)) {
      int best = -1;
      float t_near = 0.0f, t_far = 0.0f;
#pragma unroll
      for (int k = 0; k < BV_MAX_OBJECTS; ++k) {
        float tn, tf;
        if ((uint32_t)k < os.n && sphere_hit(ro, d, os.o[k].center, os.o[k].radius, &tn, &tf) &&
            (cur_k < 0 || cur_t < tn || (cur_t == tn && cur_k < k)) && (best < 0 || tn < t_near)) {
          best = k;
          t_near = tn;
          t_far = tf;
        }
      }
      if (best < 0) break;
      cur_k = best;
      cur_t = t_near;
      const ObjectView& ob = os.o[best];
      const V3 oro = xform_point(ob.m, ro);
      const V3 od = xform_dir(ob.m, d);
      V3 sd = od;
      if (!normalise(&sd)) continue;
      Colour<G, DEG> col;
      if (EVAL == BV_EVAL_SH) col.init(sd);
      Settings so = st;
      so.depth = t_far;
      WalkAcc a{0.f, 0.f, 0.f, 1.f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
      uint32_t steps, last;
      float t_end;
      walk_forward<G, U, EVAL, DEG, BOUNDED, false, true>(ob.v, so, oro, od, col, ob.start, false, &a, &steps,
                                                          &last, &t_end, nullptr, nullptr, t_near);
      const float alpha = 1.0f - a.trans;
      cr += T * a.r;
      cg += T * a.g;
      cb += T * a.b;
      T *= 1.0f - alpha;
    }
  }
  if (glane<G>() == 0) reinterpret_cast<float4*>(rgba)[i] = make_float4(cr, cg, cb, 1.0f - T);
}

}  // namespace

namespace bvh {

// Steps per ray the fused training kernel's path cache holds (256);
// BV_PATH_CACHE_STEPS overrides it (0 turns the cache off), which lets tests
// run walks that exceed the cache.
uint32_t path_cache_steps() {
  const char* env = std::getenv("BV_PATH_CACHE_STEPS");
  if (env) {
    const int v = std::atoi(env);
    if (v >= 0 && v <= 4096) return (uint32_t)v;
  }
  return 256;
}

int lanes_per_ray(int64_t rays) {
  const char* env = std::getenv("BV_LANES_PER_RAY");
  if (env) {
    const int g = std::atoi(env);
    if (g == 1 || g == 4 || g == 16 || g == 32) return g;
  }
  // Few rays cannot fill 256 CUs one lane each: spread each ray over a group.
  // Measured at 1M cells: 16 lanes win below ~128K rays (training batches),
  // 4 lanes above (camera images).
  return rays < (1 << 17) ? 16 : 4;
}

#define BV_BY_LANES(G, CALL) \
  switch (G) {               \
    case 1: CALL(1); break;  \
    case 4: CALL(4); break;  \
    case 32: CALL(32); break; \
    default: CALL(16); break; \
  }

int launch_trace_coop(bv_scene* s, const RaySource& rs, const Settings& st, const Outputs& out, int lanes,
                      hipStream_t stream) {
  const SceneView v = s->view();
  const int G = lanes ? lanes : lanes_per_ray(ray_slots(rs));
  // The pooled persistent kernel for batches of rays with their own start
  // cells (training batches, which mix cameras and so walk rays of unequal
  // length side by side): 1M random rays of 64 cameras at 1M cells trace in
  // 4.8 ms against 5.9.  Rays of one start cell (a camera's) are coherent in
  // the caller's order and walk best in lock step, as do camera sources: one
  // camera's rays in row order take 3.6 ms pooled against 3.2, its 8x8 tiles
  // 2.8 against 1.5 (the pooled kernel holds 159 VGPRs against 113 at 4 lanes,
  // 3 waves per SIMD against 4).  Its claims count 32-bit slots.
  uint32_t* next_lot = nullptr;
  int cus = 0;
  if (rs.origins != nullptr && rs.start_cells != nullptr && ray_slots(rs) < ((int64_t)1 << 31)) {
    BV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
    void* buf = nullptr;
    const int rc = stream_buffer(s, s->counter_bufs, stream, 256, &buf);
    if (rc != BV_OK) return rc;
    next_lot = static_cast<uint32_t*>(buf);
  }
#define CALL(g) launch_trace_g<g>(v, rs, st, out, next_lot, cus, stream)
  BV_BY_LANES(G, CALL)
#undef CALL
  BV_HIP(hipGetLastError());
  return BV_OK;
}

template <class RS>
int launch_record_rs(const bv_scene* s, const RS& rs, uint32_t L, float depth, float max_dt, uint32_t* cells,
                     uint32_t* next_cells, float* dts, float* mask, uint32_t* status, int lanes, hipStream_t stream) {
  const SceneView v = s->view();
  // 16 lanes per ray at every size: measured at 1M cells with 256-step rows,
  // 16 lanes beat 4 on the trainer's random pixels of 64 views (1M rows:
  // 9.4 against 10.8 ms) and on one camera's whole image in row order (786K
  // rows: 4.2 against 4.8 ms); BV_LANES_PER_RAY overrides
  const char* env = std::getenv("BV_LANES_PER_RAY");
  const int G = lanes ? lanes : (env ? lanes_per_ray(ray_slots(rs)) : 16);
#define CALL(g) launch_record_g<g>(v, rs, L, depth, max_dt, cells, next_cells, dts, mask, status, stream)
  BV_BY_LANES(G, CALL)
#undef CALL
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int launch_record_coop(const bv_scene* s, const RaySource& rs, uint32_t L, float depth, float max_dt, uint32_t* cells,
                       uint32_t* next_cells, float* dts, float* mask, uint32_t* status, int lanes,
                       hipStream_t stream) {
  return launch_record_rs(s, rs, L, depth, max_dt, cells, next_cells, dts, mask, status, lanes, stream);
}

int launch_record_coop(const bv_scene* s, const ViewRays& rs, uint32_t L, float depth, float max_dt, uint32_t* cells,
                       uint32_t* next_cells, float* dts, float* mask, uint32_t* status, int lanes,
                       hipStream_t stream) {
  return launch_record_rs(s, rs, L, depth, max_dt, cells, next_cells, dts, mask, status, lanes, stream);
}

// Launch plan of one fused training call: lanes per ray, the persistent grid,
// the path cache length and whether density / position gradients go through
// the packed per-cell table.  `budget` bounds the path cache's bytes.
struct TrainPlan {
  int G = 16;
  int64_t blocks = 1;
  size_t groups = 0;
  uint32_t cap = 0;
  bool pack = false;
  size_t off_pack = 0;  // byte offset of the packed table in the work buffer
  size_t bytes = 0;     // work buffer bytes: ray counter, path cache, packed table
};

constexpr size_t kCounterBytes = 256;
constexpr size_t kPathStepBytes = 4 * sizeof(float4);
// the path cache gets up to 256 MiB of a per-stream buffer
constexpr size_t kPathBudget = size_t(256) << 20;

static int plan_train(const bv_scene* s, const RaySource& rs, int lanes, size_t budget, bool allow_pack, bool adj,
                      TrainPlan* p) {
  // Training rays are incoherent: 16 lanes per ray beat 4 at every batch size
  // measured (1M cells, 4K to 1M rays per call; 32 lanes tie), unlike camera
  // images, whose neighbouring rays share cells.
  const char* env = std::getenv("BV_LANES_PER_RAY");
  p->G = lanes ? lanes : (env ? lanes_per_ray(rs.count) : 16);
  // Grid: the blocks the rays need, at most as many as the device holds at
  // once (the kernel loops over the rest).
  int per_cu = 1;
#define CALL(gg) per_cu = train_blocks_per_cu_g<gg>(s->sh_degree, adj)
  BV_BY_LANES(p->G, CALL)
#undef CALL
  int cus = 0;
  BV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device));
  const int64_t need = (rs.count * p->G + 255) / 256;
  p->blocks = std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)std::max(cus, 1) * per_cu));
  p->groups = (size_t)p->blocks * (256 / p->G);
  // Path cache: up to path_cache_steps() steps per group (64 B each) within
  // the budget; rays longer than the cache are replayed by walking them again.
  p->cap = (uint32_t)std::min<size_t>(path_cache_steps(), budget / (kPathStepBytes * p->groups));
  if (p->cap < 8) p->cap = 0;
  // then the packed density / position gradients, one float4 per cell, when
  // the call is large enough (about 37 segments a ray) for one fewer memory
  // atomic per segment to outweigh clearing and adding the whole table
  p->pack = allow_pack && rs.count * 16 >= s->n;
  p->off_pack = (kCounterBytes + kPathStepBytes * p->cap * p->groups + 255) & ~size_t(255);
  p->bytes = p->off_pack + (p->pack ? sizeof(float4) * (size_t)s->n : 0);
  return BV_OK;
}

// One fused training launch over a work buffer laid out by plan_train: the
// ray counter (reset by a kernel), the path cache, then the packed table.
static int run_train(const bv_scene* s, const RaySource& rs, const Settings& st, const RayTerms& t, bool adj,
                     const GradPtrs& g, const TrainPlan& p, void* buf, hipStream_t stream, uint32_t* path = nullptr,
                     uint32_t path_len = 0) {
  const SceneView v = s->view();
  uint32_t* next_ray = reinterpret_cast<uint32_t*>(buf);
  const PathCache pc{p.cap ? reinterpret_cast<float4*>(static_cast<char*>(buf) + kCounterBytes) : nullptr, p.cap};
  GradPtrs gp = g;
  const unsigned cgrid = (unsigned)std::max<int64_t>(1, (s->n + 255) / 256);
  if (p.pack) {
    gp.packed = reinterpret_cast<float4*>(static_cast<char*>(buf) + p.off_pack);
    hipLaunchKernelGGL(k_zero_packed, dim3(cgrid), dim3(256), 0, stream, (int64_t)s->n, gp.packed);
  }
#define CALL(gg) \
  launch_train_g<gg>(v, rs, st, t, adj, gp, pc, next_ray, (unsigned)p.blocks, path, path_len, stream)
  BV_BY_LANES(p.G, CALL)
#undef CALL
  if (p.pack)
    hipLaunchKernelGGL(k_unpack_grads, dim3(cgrid), dim3(256), 0, stream, (int64_t)s->n, (const float4*)gp.packed,
                       g.d_density, g.d_xyz);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int launch_train_l2_coop(bv_scene* s, const RaySource& rs, const Settings& st, const float* target,
                         float* rgba_out, float* loss, const GradPtrs& g, int lanes, hipStream_t stream,
                         uint32_t* path, uint32_t path_len) {
  if (rs.count <= 0) return BV_OK;
  TrainPlan p;
  int rc = plan_train(s, rs, lanes, kPathBudget, true, false, &p);
  if (rc != BV_OK) return rc;
  void* buf = nullptr;
  rc = stream_buffer(s, s->path_bufs, stream, p.bytes, &buf);
  if (rc != BV_OK) return rc;
  return run_train(s, rs, st, RayTerms{target, rgba_out, loss, nullptr, nullptr}, false, g, p, buf, stream, path,
                   path_len);
}

// The fast adjoint runs the fused kernel's machinery on the caller's forward
// colours and loss gradients: a persistent launch whose groups walk each ray
// again only to fill the path cache, then replay it step-parallel (the walk
// again, sequentially, for rays longer than the cache), first steps summed
// per cell in LDS.
int launch_adjoint_coop(bv_scene* s, const RaySource& rs, const Settings& st, const float* rgba_fwd,
                        const float* d_rgba, const GradPtrs& g, int lanes, hipStream_t stream) {
  if (rs.count <= 0) return BV_OK;
  TrainPlan p;
  int rc = plan_train(s, rs, lanes, kPathBudget, true, true, &p);
  if (rc != BV_OK) return rc;
  void* buf = nullptr;
  rc = stream_buffer(s, s->path_bufs, stream, p.bytes, &buf);
  if (rc != BV_OK) return rc;
  return run_train(s, rs, st, RayTerms{nullptr, nullptr, nullptr, rgba_fwd, d_rgba}, true, g, p, buf, stream);
}

}  // namespace bvh

// ---------------------------------------------------------------- allocation-free training
// Workspace of bv_train_step_l2_noalloc: the ray counter, then the path cache.
// Each device that runs the library holds its own copy until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host the
// name is only a stand-in): it uses the device address from
// hipGetSymbolAddress, kept per device.
static __device__ __align__(256) unsigned char scratch[bvh::kNoAllocScratchBytes];

namespace bvh {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The device address of this device's `scratch`, looked up once per device
// with that device current (it is on the calling thread).
static int noalloc_scratch(void** out) {
  int dev = 0;
  BV_HIP(hipGetDevice(&dev));
  BV_REQUIRE(dev >= 0 && dev < kMaxDevices, BV_ERR_UNSUPPORTED, "device index beyond the workspace table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    BV_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    BV_REQUIRE(ptr != nullptr, BV_ERR_HIP, "hipGetSymbolAddress returned a null workspace");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return BV_OK;
}

int launch_train_l2_noalloc(const bv_scene* s, const RaySource& rs, const Settings& st, const float* target,
                            float* rgba_out, float* loss, const GradPtrs& g, hipStream_t stream) {
  if (rs.count <= 0) return BV_OK;
  void* buf = nullptr;
  int rc = noalloc_scratch(&buf);
  if (rc != BV_OK) return rc;
  // no packed table: density and position gradients go straight to the
  // caller's arrays, so the workspace holds only the counter and the cache
  TrainPlan p;
  rc = plan_train(s, rs, 0, kNoAllocScratchBytes - kCounterBytes, false, false, &p);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(p.bytes <= kNoAllocScratchBytes, BV_ERR_CAPACITY, "training plan exceeds the fixed workspace");
  return run_train(s, rs, st, RayTerms{target, rgba_out, loss, nullptr, nullptr}, false, g, p, buf, stream);
}

int launch_relight_coop(const bv_scene* s, const RaySource& rs, const Settings& st, const bv_lighting& lit,
                        float* rgba_out, int lanes, hipStream_t stream) {
  const SceneView v = s->view();
  const int G = lanes ? lanes : lanes_per_ray(ray_slots(rs));
#define CALL(gg) launch_relight_g<gg>(v, rs, st, lit, rgba_out, stream)
  BV_BY_LANES(G, CALL)
#undef CALL
  BV_HIP(hipGetLastError());
  return BV_OK;
}

// The serving kernel of one (mode, SH degree, boundedness) variant: its
// kernel and the blocks of it one CU holds at once.
template <int M, int D, bool B>
static const void* serve_kernel() {
  return reinterpret_cast<const void*>(&k_serve<M, D, B>);
}

static const void* serve_kernel_of(int mode, int sh_degree, bool bounded) {
  const int deg = sh_degree < 3 ? sh_degree : 3;
#define BV_SK(M)                                                                        \
  switch (deg) {                                                                        \
    case 0: return bounded ? serve_kernel<M, 0, true>() : serve_kernel<M, 0, false>();  \
    case 1: return bounded ? serve_kernel<M, 1, true>() : serve_kernel<M, 1, false>();  \
    case 2: return bounded ? serve_kernel<M, 2, true>() : serve_kernel<M, 2, false>();  \
    default: return bounded ? serve_kernel<M, 3, true>() : serve_kernel<M, 3, false>(); \
  }
  if (mode == BV_EVAL_RELIGHT) {
    BV_SK(BV_EVAL_RELIGHT)
  } else {
    BV_SK(BV_EVAL_SH)
  }
#undef BV_SK
}

int serve_blocks_per_cu(int mode, int sh_degree, bool bounded) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, serve_kernel_of(mode, sh_degree, bounded), 256, 0) !=
          hipSuccess ||
      b < 1)
    b = 1;
  return b;
}

int launch_serve(int mode, int sh_degree, bool bounded, const bvk::ServeQueue& q, uint32_t blocks,
                 hipStream_t stream) {
  void* args[] = {const_cast<bvk::ServeQueue*>(&q)};
  BV_HIP(hipLaunchKernel(serve_kernel_of(mode, sh_degree, bounded), dim3(blocks), dim3(256), args, 0, stream));
  return BV_OK;
}

int launch_trace_objects(const bvk::ObjectSet& os, int32_t sh_degree, bool bounded, const float* origins,
                         const float* dirs, int64_t count, const bvk::Settings& st, float* rgba, hipStream_t stream) {
  if (count <= 0) return BV_OK;
  // 4 lanes per ray, the single-scene walk's choice for large calls (one
  // 1024x768 camera at 1M cells, one identity object: 127.7 Mrays/s with one
  // lane per ray against 344.8 for the 4-lane walk)
  constexpr int G = 4;
  constexpr int U = unroll_for(G);
  const int64_t threads = count * G;
  const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
#define BV_O(EV, D)                                                                                             \
  if (bounded)                                                                                                  \
    hipLaunchKernelGGL((k_trace_objects<G, U, EV, D, true>), grid, block, 0, stream, os, origins, dirs, count, st, rgba); \
  else                                                                                                          \
    hipLaunchKernelGGL((k_trace_objects<G, U, EV, D, false>), grid, block, 0, stream, os, origins, dirs, count, st, rgba);
#define BV_OD(EV)               \
  switch (sh_degree) {          \
    case 0: BV_O(EV, 0) break;  \
    case 1: BV_O(EV, 1) break;  \
    case 2: BV_O(EV, 2) break;  \
    default: BV_O(EV, 3) break; \
  }
  if (st.eval_mode == BV_EVAL_CONSTANT) {
    BV_OD(BV_EVAL_CONSTANT)
  } else {
    BV_OD(BV_EVAL_SH)
  }
#undef BV_OD
#undef BV_O
  BV_HIP(hipGetLastError());
  return BV_OK;
}

}  // namespace bvh
