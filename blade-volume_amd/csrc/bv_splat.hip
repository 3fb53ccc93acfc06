// PowerFoam splat tracer: every support sphere the ray hits is clipped by the
// radical planes of its adjacency row (the Cech graph of the supports), and the
// surviving intervals are composited in (near, cell) order.
//
// Reference semantics: trace_powerfoam_splats / sorted_powerfoam_splat_intervals /
// powerfoam_splat_interval (blade-volume/src/trace.rs:1331-1546), which the
// reference GPU path reproduces with a candidate gather, a per-candidate clip
// and an ordered selection (shaders/radfoam_record_paths.wgsl:579-1218) followed
// by integrate_powerfoam_splats (shaders/powerfoam_splat.wgsl:246-289).
//
// Design on gfx950:
//   * Candidates come from a linear BVH over the support spheres, built on the
//     device once per scene geometry (Morton codes of the sites, radix sort,
//     Karras split search, bottom-up box fit).  Nodes hold both children's boxes
//     (64 B), so one node fetch tests two boxes.  Leaves are sites in Morton
//     order with their 32-byte cell record copied alongside, so a leaf test is
//     one 32-byte load.
//   * The walk over the sorted interval list is a k-buffer with a cursor: a
//     traversal collects the K smallest (near, cell) keys beyond the cursor,
//     front-to-back traversal and box pruning against the K-th key keep it
//     local, the K intervals are composited, and the next traversal resumes
//     after the last key.  Transmittance early-out ends the ray as soon as the
//     reference's loop would stop, so only the front of the sorted list is ever
//     built.  The result does not depend on K.
//   * The clip of one candidate reads the cell's 32-byte edge records
//     (neighbour position, radius and the precomputed radical-plane shift
//     0.5 + 0.5 (r_i^2 - r_j^2) / max(|p_j - p_i|^2, 1e-20)) and forms
//     p_j - p_i, the face origin and both dot products with the reference's
//     operation order, so every interval bound is bit-identical to the
//     reference's FP32 value.  Because the entry bound only rises and the exit
//     bound only falls while planes are applied, the loop stops as soon as the
//     interval is empty, lies before the cursor or beyond the K-th key: those
//     exits cannot change the result.
//   * Boxes are padded (several ulp of the site scale) and pruning thresholds
//     carry a slack, so a sphere is never culled by rounding: the exact sphere
//     and plane arithmetic alone decides every interval.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "bv_bvh.h"
#include "bv_host.h"

namespace bvk {

// Leaves in Morton order: the cell record of each site and its index.
__global__ void k_leaves(int64_t n, const uint32_t* __restrict__ perm, const CellRec* __restrict__ cells,
                         CellRec* leaf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  leaf[i] = cells[perm[i]];
}

// Support-sphere boxes padded by several ulp of the site scale.
__global__ void k_sphere_boxes(int64_t n, const float4* __restrict__ pts, float4* lo, float4* hi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sphere_box(pts[i], &lo[i], &hi[i]);
}

struct SplatRecord {
  uint32_t* cells;
  float* nears;
  float* dts;
  uint32_t* count;
  uint32_t L;
};

// The PowerFoam recorder's rows with Jacobian streams (PathRecordBuffers with
// jacobian_mode 1, gpu/path_record.rs:614-670): [P, L] cell, next, previous,
// dt, mask, three float4 Jacobian rows and the reference tangent; status [P].
struct SplatJacRecord {
  uint32_t* cells;
  uint32_t* next;
  uint32_t* prev;
  float* dts;
  float* mask;
  float4* jp;
  float4* jc;
  float4* jn;
  float* tangent;
  uint32_t* status;
  uint32_t L;
  float max_dt;
};

struct SplatCounters {
  uint32_t nodes, leaves, planes, rounds;
};

// The intervals of one ray in (near, cell) order, handed to visit(cell, near,
// dt) until it returns false or the list ends.  d is unit length.
//
// One flat loop advances the lane by one unit of work per iteration: pop a
// reference (and visit an inner node), test one support sphere, or apply four
// radical planes of the candidate being clipped.  Lanes of a wave thus stay in
// step however unevenly their traversals, leaf ranges and rows (the Cech rows
// of large supports are long) are distributed.  A round ends when the stack is
// empty; its buffered intervals are final and go to visit in order, all lanes
// of the wave together once every live lane has ended its round (visits are
// the costly part of shading, and lanes visiting one at a time leave the rest
// of the wave idle).
//
// Per candidate (powerfoam_splat_interval + the bounded branch of
// path_interval_jacobian): camera exclusion at 4 r, sphere roots -b -+ sqrt(disc)
// inside (0, depth), then every plane of the row: entry faces raise fn, exit
// faces lower ff, a parallel plane with the ray on its far side empties the
// interval.  The exits taken while planes are applied (interval already empty,
// wholly before the cursor, or past the current key limit) cannot change the
// result, because fn only rises and ff only falls.
//
// Key limit.  Every visit here stops before an interval once `stop_count`
// intervals were visited or the transmittance is <= `stop_T` (alpha = 1 -
// exp(-density dt) per interval with density > 1e-6).  While the buffer is not
// full the stream keeps the product of its entries' factors (1 - alpha, formed
// as the visit forms it), i.e. the transmittance after its last entry.  Once
// that product times the transmittance at the round's start is <= stop_T
// (less a relative margin far above the rounding of the two products), or the
// buffer holds as many entries as visits are left, the visits end at or before
// the buffer's last key: no interval past it can be visited, so that key
// bounds the window instead of the K-th key of a full buffer.  An eviction
// from a full buffer drops the bound back to the K-th key.
template <class Visit>
__device__ __forceinline__ void splat_stream(const SceneView& s, const BvhView& b, V3 o, V3 d, float depth,
                                             uint32_t (*stk)[kSplatBlock], SplatCounters& ctr, float stop_T,
                                             uint32_t stop_count, Visit&& visit) {
  const int64_t n = s.n;
  if (n <= 0) return;
  const uint32_t lane = threadIdx.x;
  const uint32_t inner = (uint32_t)(n > 1 ? n - 1 : 1);
  const V3 inv = mk(1.0f / (d.x != 0.0f ? d.x : 1e-30f), 1.0f / (d.y != 0.0f ? d.y : 1e-30f),
                    1.0f / (d.z != 0.0f ? d.z : 1e-30f));
  const float depth_hi = depth + slack(depth);
  uint64_t cursor = 0;      // intervals already handed out have keys <= cursor
  float ct = -INFINITY;     // the cursor's near distance
  float lo_base = -slack(0.0f);
  KBuf kb;
  kb.clear(false);
  const float stop_bound = stop_T * (1.0f - 1e-4f);
  float t_round = 1.0f;     // transmittance before the round's first entry (an upper bound)
  float prod = 1.0f;        // product of the buffered entries' factors while the buffer is not full
  uint64_t last = 0;        // largest buffered key
  bool ends = false;        // the visits end at or before `last`
  uint32_t visited = 0;     // intervals handed to visit in earlier rounds
  int sp = 0;
  stk[sp++][lane] = 0u;  // the root
  uint32_t lpos = 0, lend = 0;  // leaves still to test in the current range
  bool clipping = false;
  // candidate being clipped
  uint32_t ccell = 0, e = 0, e_end = 0;
  V3 cp = mk(0.f, 0.f, 0.f);
  float cden = 0.f, sn = 0.f, sf = 0.f, fn = 0.f, ff = 0.f;
  // Round complete: the buffered intervals are final and go to visit in
  // order.  False once the stream has ended for this lane.
  auto finish_round = [&]() -> bool {
    const uint32_t cnt = kb.count;
    const bool was_full = kb.full();
    const uint64_t kth = kb.kth();
    ctr.rounds += 1;
    for (uint32_t i = 0; i < cnt; ++i) {
      const uint64_t k0 = kb.k[0];
      const float dt0 = kb.v[0];
      kb.pop_front();
      if (!visit(key_id(k0), key_t(k0), dt0)) return false;
    }
    if (!was_full) return false;
    // a full round: every entry was visited; `prod` is stale after an
    // eviction, so the transmittance bound is kept as is
    visited += cnt;
    cursor = kth;
    ct = key_t(kth);
    lo_base = ct - slack(ct);
    kb.clear(false);
    prod = 1.0f;
    last = 0;
    ends = false;
    stk[sp++][lane] = 0u;
    return true;
  };
  // A lane whose round is complete waits until every live lane of the wave
  // has completed its round; the wave then visits together, so the visits
  // (shading every light, evaluating SH rows) run on all lanes at once
  // instead of on the few whose rounds happened to end in the same
  // iteration.
  bool waiting = false;
  // One back-edge only (no `continue`): a loop with several back-edges is
  // split by the compiler into nested loops, which serialises the node, leaf
  // and clipping steps of the lanes of a wave.
  for (;;) {
    if (__ballot(!waiting) == 0ull) {
      if (!finish_round()) return;
      waiting = false;
    }
    if (!clipping && !waiting) {
      if (lpos == lend) {
        if (sp == 0) {
          waiting = true;
        } else {
          const uint32_t ref = stk[--sp][lane];
          if (ref & kLeaf) {
            lpos = ref & kLeafPos;
            lend = lpos + ((ref >> 29) & 3u) + 1u;
            if (n == 1 || (uint64_t)lend > (uint64_t)n) lend = lpos + (lpos < (uint64_t)n ? 1u : 0u);
          } else if (ref < inner) {
            const BvhNode nd = b.nodes[ref];
            ++ctr.nodes;
            const float kt = key_t(ends ? last : kb.kth());
            const float hi_t = fminf(depth_hi, kt + slack(kt));
            float tl, tr;
            const bool hl = box_hit(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, o, inv, lo_base, hi_t, &tl);
            // a single-site tree repeats its leaf in the right slot
            const bool hr =
                n != 1 && box_hit(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, o, inv, lo_base, hi_t, &tr);
            // push the far child first so the near one is taken next
            const bool near_right = hl && hr && tr < tl;
            const uint32_t first = near_right ? nd.d.x : nd.d.y;
            const uint32_t second = near_right ? nd.d.y : nd.d.x;
            const bool first_hit = near_right ? hl : hr;
            const bool second_hit = near_right ? hr : hl;
            if (first_hit && sp < (int)b.stack_rows) stk[sp++][lane] = first;
            if (second_hit && sp < (int)b.stack_rows) stk[sp++][lane] = second;
          }
        }
      }
      // Opaque to the optimiser: otherwise it threads the node path straight
      // back to the loop header, which recreates a nested node loop.
      __asm__ volatile("" : "+v"(lpos));
      if (lpos != lend) {
        // sphere test of the next leaf
        const CellRec cr = b.leaf[lpos];
        ccell = b.cell[lpos];
        ++lpos;
        ++ctr.leaves;
        cp = mk(cr.x, cr.y, cr.z);
        cden = cr.density;
        const float r = cr.radius;
        const V3 cam = sub(cp, o);
        const float ex = 4.0f * r;
        if (!(dot(cam, cam) < ex * ex)) {
          const V3 oc = sub(o, cp);
          const float bq = dot(oc, d);
          const float disc = bq * bq - (dot(oc, oc) - r * r);
          if (disc > 0.0f && isfinite(disc)) {
            const float root = sqrtf(disc);
            sn = -bq - root;
            sf = -bq + root;
            const float kt = key_t(ends ? last : kb.kth());
            if (!(sf <= 0.0f || sn >= depth) && !(sf < ct) && !(sn > kt)) {
              fn = 0.0f;
              ff = depth;
              e = cr.begin;
              e_end = cr.end;
              clipping = true;
            }
          }
        }
      }
    }
    if (clipping) {
      // four planes of the candidate's row per iteration, loads issued together
      EdgeRec er[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) er[q] = s.erec[e + q < e_end ? e + q : (e_end > 0 ? e_end - 1 : 0)];
      const uint64_t kth = ends ? last : kb.kth();
      const float kt = key_t(kth);
      bool dead = false;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!dead && e + q < e_end) {
          ++ctr.planes;
          const V3 nrm = mk(er[q].x - cp.x, er[q].y - cp.y, er[q].z - cp.z);
          const V3 fo = add(cp, scl(er[q].shift, nrm));
          const float num = dot(sub(fo, o), nrm);
          const float den = dot(d, nrm);
          if (den > 1e-20f) {
            const float t = num / den;
            if (t < ff) ff = t;
          } else if (den < -1e-20f) {
            const float t = num / den;
            if (t > fn) fn = t;
          } else if (num < 0.0f) {
            dead = true;
          }
          const float cs = sn > fn ? sn : fn;
          const float ce = sf < ff ? sf : ff;
          if (ce <= cs || ce <= ct || cs > kt) dead = true;
        }
      }
      e += 4;
      if (dead) {
        clipping = false;
      } else if (e >= e_end) {
        clipping = false;
        const float start = sn > fn ? sn : fn;
        const float end = sf < ff ? sf : ff;
        if (end > start && isfinite(start) && isfinite(end)) {
          const uint64_t key = sort_key(start, ccell);
          if (key > cursor && key < kth) {
            const float dt = end - start;
            if (kb.full()) {
              ends = false;  // the evicted entry's factor is unknown
            } else {
              const float alpha = 1.0f - expf(-cden * dt);
              prod *= cden > 1e-6f ? 1.0f - alpha : 1.0f;
              last = key > last ? key : last;
              ends = t_round * prod <= stop_bound || visited + kb.count + 1u >= stop_count;
            }
            kb.insert(key, dt);
          }
        }
      }
    }
  }
}

// The derivatives of one recorded interval's length (path_interval_jacobian,
// trace.rs:1199-1318, bounded branch; interval_differential,
// radfoam_record_paths.wgsl:291-430).  The visit hands over (cell, near, dt);
// the cell's row is clipped again with the stream's arithmetic, now keeping
// which plane set the entry (the previous cell) and the exit (the next cell),
// so near and dt come out bit for bit as the stream found them.
struct SplatJacRow {
  uint32_t prev, next;
  float dt;
  float4 jp, jc, jn;
  float tangent;
};

// face_intersection_jacobians (trace.rs:1077-1097): d t / d (xyz, r) of the
// current and adjacent sites for the radical plane crossed at t.
__device__ __forceinline__ void face_jac4(V3 o, V3 d, V3 cur, float rc, V3 adj, float ra, float t, float4* jc,
                                          float4* ja) {
  const V3 nrm = sub(adj, cur);
  const float den = dot(d, nrm);
  if (!isfinite(den) || fabsf(den) <= 1e-20f) {
    *jc = *ja = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const V3 a = add(sub(o, cur), scl(t, d));
  const V3 b = sub(sub(adj, o), scl(t, d));
  *jc = make_float4(a.x / den, a.y / den, a.z / den, rc / den);
  *ja = make_float4(b.x / den, b.y / den, b.z / den, -ra / den);
}

__device__ __forceinline__ float dot4(float4 j, V3 g, float w) { return ((j.x * g.x + j.y * g.y) + j.z * g.z) + j.w * w; }

__device__ __forceinline__ void splat_jacobian(const SceneView& s, V3 o, V3 d, float depth, uint32_t cell,
                                               float max_dt, SplatJacRow* out) {
  const CellRec cr = s.cells[cell];
  const V3 cp = mk(cr.x, cr.y, cr.z);
  const float r = cr.radius;
  // sphere_intersection_jacobians (trace.rs:1101-1124)
  const V3 oc = sub(o, cp);
  const float bq = dot(oc, d);
  const float disc = bq * bq - (dot(oc, oc) - r * r);
  const float root = sqrtf(disc);
  const float sn = -bq - root, sf = -bq + root;
  const V3 perp = sub(oc, scl(bq, d));
  const V3 rdc = mk(perp.x / root, perp.y / root, perp.z / root);
  const float rdr = r / root;
  // the entry and exit planes of the row (powerfoam_splat_interval, trace.rs:1331-1389)
  float fn = 0.0f, ff = depth;
  uint32_t prev = cell, next = cell;
  V3 pp = cp, pn = cp;
  float rp = r, rn = r;
  for (uint32_t e = cr.begin; e < cr.end; ++e) {
    const EdgeRec er = s.erec[e];
    const V3 nrm = mk(er.x - cp.x, er.y - cp.y, er.z - cp.z);
    const V3 fo = add(cp, scl(er.shift, nrm));
    const float num = dot(sub(fo, o), nrm);
    const float den = dot(d, nrm);
    if (den > 1e-20f) {
      const float t = num / den;
      if (t < ff) {
        ff = t;
        next = er.nb;
        pn = mk(er.x, er.y, er.z);
        rn = er.radius;
      }
    } else if (den < -1e-20f) {
      const float t = num / den;
      if (t > fn) {
        fn = t;
        prev = er.nb;
        pp = mk(er.x, er.y, er.z);
        rp = er.radius;
      }
    }
  }
  const bool start_sphere = sn > fn, end_sphere = sf < ff;
  const float start = start_sphere ? sn : fn;
  const float end = end_sphere ? sf : ff;
  float4 jp = make_float4(0.f, 0.f, 0.f, 0.f), jc = jp, jn = jp;
  if (start_sphere) {
    jc = make_float4(-(d.x - rdc.x), -(d.y - rdc.y), -(d.z - rdc.z), rdr);
  } else if (prev != cell) {
    float4 qc, qp;
    face_jac4(o, d, cp, r, pp, rp, fn, &qc, &qp);
    jc = make_float4(-qc.x, -qc.y, -qc.z, -qc.w);
    jp = make_float4(-qp.x, -qp.y, -qp.z, -qp.w);
  }
  if (end_sphere) {
    jc = make_float4(jc.x + (d.x + rdc.x), jc.y + (d.y + rdc.y), jc.z + (d.z + rdc.z), jc.w + rdr);
  } else if (next != cell) {
    float4 qc, qn;
    face_jac4(o, d, cp, r, pn, rn, ff, &qc, &qn);
    jc = make_float4(jc.x + qc.x, jc.y + qc.y, jc.z + qc.z, jc.w + qc.w);
    jn = qn;
  }
  float dt = end - start;
  if (dt > max_dt) {
    // the clamp is locally constant on its saturated branch
    dt = max_dt;
    jp = jc = jn = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  out->prev = prev;
  out->next = next;
  out->dt = dt;
  out->jp = jp;
  out->jc = jc;
  out->jn = jn;
  out->tangent = dot4(jp, sub(pp, o), rp) + dot4(jc, sub(cp, o), r) + dot4(jn, sub(pn, o), rn);
}

// Front-to-back compositing of trace_powerfoam_splats (trace.rs:1497-1536).
template <int EVAL, int DEG>
struct SplatTrace {
  WalkAcc a;
  uint32_t steps, last;
  float t_end;

  __device__ __forceinline__ void init(uint32_t start) {
    a = WalkAcc{0.f, 0.f, 0.f, 1.0f, 0.f, 0.f, BV_NO_CELL, 0u, 0u};
    steps = 0;
    last = start;
    t_end = 0.0f;
  }
  __device__ __forceinline__ bool step(const SceneView& s, const Settings& st, V3 d, uint32_t cell, float near_t,
                                       float dt) {
    constexpr int K = (DEG + 1) * (DEG + 1);
    if (steps == st.max_steps || a.trans <= st.weight_threshold) return false;
    steps += 1;
    last = cell;
    t_end = near_t + dt;
    const float* row = s.attrs + (int64_t)cell * s.attr_dim;
    const float sd = row[3 * K];
    a.bytes += 4u;
    if (sd <= 1e-6f) return true;
    const float alpha = 1.0f - expf(-sd * dt);
    const float w = a.trans * alpha;
    if (EVAL == BV_EVAL_CONSTANT) {
      a.r += w * st.const_rgb[0];
      a.g += w * st.const_rgb[1];
      a.b += w * st.const_rgb[2];
    } else if (EVAL == BV_EVAL_SH) {
      const V3 col = clamp0(eval_sh_row<DEG>(row, d));
      a.r += w * col.x;
      a.g += w * col.y;
      a.b += w * col.z;
      a.bytes += 12u * K;
    }
    if (w > a.peak) {
      a.peak = w;
      a.depth_mode = near_t + 0.5f * dt;
      a.peak_point = cell;
    }
    a.trans *= 1.0f - alpha;
    a.segments += 1;
    return true;
  }
};

// Relit splats: each composited interval is shaded by every light at its entry
// point x = o + near d, normal x - site, view -d and the cell's material; light
// l's sum is scaled by the transmittance of a splat opacity trace from the
// depth point toward the light when visibility is on.
template <int DEG>
__device__ float4 splat_relight(const SceneView& s, const BvhView& b, const Settings& st, const bv_lighting& lit, V3 o,
                                V3 d, uint32_t (*stk)[kSplatBlock], float* lsum, SplatCounters& ctr) {
  constexpr int K = (DEG + 1) * (DEG + 1);
  const int nl = lit.count < BV_MAX_LIGHTS ? (int)lit.count : BV_MAX_LIGHTS;
  // Per-light colour sums live in LDS (lsum [3 BV_MAX_LIGHTS][kSplatBlock],
  // this lane's column): the light loop can then index them at run time
  // without the 24 registers (and the unrolled shading) that kept the kernel
  // at two waves per SIMD.
  float* sums = lsum + threadIdx.x;
  for (int q = 0; q < 3 * nl; ++q) sums[q * kSplatBlock] = 0.f;
  const V3 view = mk(-d.x, -d.y, -d.z);
  float trans = 1.0f, peak = 0.0f, depth_mode = 0.0f;
  uint32_t peak_point = BV_NO_CELL, steps = 0;
  splat_stream(s, b, o, d, st.depth, stk, ctr, st.weight_threshold, st.max_steps,
                     [&](uint32_t cell, float near_t, float dt) {
    if (steps == st.max_steps || trans <= st.weight_threshold) return false;
    steps += 1;
    const float sd = s.attrs[(int64_t)cell * s.attr_dim + 3 * K];
    if (sd <= 1e-6f) return true;
    const float alpha = 1.0f - expf(-sd * dt);
    const float w = trans * alpha;
    const float4 p4 = s.pts[cell];
    const V3 x = add(o, scl(near_t, d));
    const V3 nrm = sub(x, mk(p4.x, p4.y, p4.z));
    const float4 mat = s.material[cell];
#pragma unroll 1
    for (int l = 0; l < nl; ++l) {
      const V3 col = shade_point_light(lit.lights[l], x, nrm, view, mat, lit.specular_f0);
      sums[(3 * l) * kSplatBlock] += w * col.x;
      sums[(3 * l + 1) * kSplatBlock] += w * col.y;
      sums[(3 * l + 2) * kSplatBlock] += w * col.z;
    }
    if (w > peak) {
      peak = w;
      depth_mode = near_t + 0.5f * dt;
      peak_point = cell;
    }
    trans *= 1.0f - alpha;
    return true;
  });
  float4 rgba = make_float4(0.f, 0.f, 0.f, 1.0f - trans);
  const bool shadows = peak_point != BV_NO_CELL;
  const V3 p = add(o, scl(depth_mode, d));
  for (int l = 0; l < nl; ++l) {
    float vis = 1.0f;
    // a visibility ray for every light (lit.visibility) or for the lights that ask for one
    if (shadows && (lit.visibility != 0 || lit.lights[l].visibility != 0)) {
      V3 to = sub(mk(lit.lights[l].position[0], lit.lights[l].position[1], lit.lights[l].position[2]), p);
      const float dist = sqrtf(dot(to, to));
      if (dist > 0.0f && normalise(&to)) {
        float sh_trans = 1.0f;
        uint32_t sh_steps = 0;
        splat_stream(s, b, p, to, dist, stk, ctr, st.weight_threshold, st.max_steps,
                     [&](uint32_t cell, float, float dt) {
          if (sh_steps == st.max_steps || sh_trans <= st.weight_threshold) return false;
          sh_steps += 1;
          const float sd = s.attrs[(int64_t)cell * s.attr_dim + 3 * K];
          if (sd > 1e-6f) {
            const float alpha = 1.0f - expf(-sd * dt);
            sh_trans *= 1.0f - alpha;
          }
          return true;
        });
        vis = sh_trans;
      }
    }
    rgba.x += vis * sums[(3 * l) * kSplatBlock];
    rgba.y += vis * sums[(3 * l + 1) * kSplatBlock];
    rgba.z += vis * sums[(3 * l + 2) * kSplatBlock];
  }
  return rgba;
}

// MODE 0: trace (trace_powerfoam_splats); MODE 1: record the first L intervals.
template <int MODE, int EVAL, int DEG, class RS>
__global__ __launch_bounds__(kSplatBlock) void k_splat(SceneView s, BvhView b, RS rs, Settings st,
                                                       Outputs out, SplatRecord rec, SplatJacRecord jr) {
  extern __shared__ uint32_t lds_stack[];
  auto stk = reinterpret_cast<uint32_t (*)[kSplatBlock]>(lds_stack);
  const int64_t slot = (int64_t)blockIdx.x * kSplatBlock + threadIdx.x;
  if (slot >= ray_slots(rs)) return;
  V3 o, d;
  uint32_t start;
  int64_t idx;
  if (!fetch_ray(rs, slot, &o, &d, &start, &idx)) return;
  if (start == BV_NO_CELL) start = 0;
  SplatCounters ctr{0u, 0u, 0u, 0u};
  const bool ok = normalise(&d);
  if (MODE == 1) {
    uint32_t written = 0;
    if (ok)
      splat_stream(s, b, o, d, st.depth, stk, ctr, -1.0f, rec.L, [&](uint32_t cell, float near_t, float dt) {
        if (written == rec.L) return false;
        const int64_t row = idx * (int64_t)rec.L + written;
        rec.cells[row] = cell;
        rec.nears[row] = near_t;
        rec.dts[row] = dt;
        ++written;
        return true;
      });
    rec.count[idx] = written;
    return;
  }
  if (MODE == 2) {
    // one interval past L tells whether the row was truncated
    uint32_t written = 0;
    bool truncated = false;
    if (ok)
      splat_stream(s, b, o, d, st.depth, stk, ctr, -1.0f, jr.L + 1u, [&](uint32_t cell, float near_t, float dt) {
        if (written == jr.L) {
          truncated = true;
          return false;
        }
        SplatJacRow j;
        splat_jacobian(s, o, d, st.depth, cell, jr.max_dt, &j);
        const int64_t row = idx * (int64_t)jr.L + written;
        jr.cells[row] = cell;
        jr.next[row] = j.next;
        jr.prev[row] = j.prev;
        jr.dts[row] = j.dt;
        if (jr.mask) jr.mask[row] = 1.0f;
        jr.jp[row] = j.jp;
        jr.jc[row] = j.jc;
        jr.jn[row] = j.jn;
        jr.tangent[row] = j.tangent;
        ++written;
        return true;
      });
    jr.status[idx] = written | (truncated ? BV_PATH_TRUNCATED : 0u);
    return;
  }
  SplatTrace<EVAL, DEG> tr;
  tr.init(start);
  if (ok)
    splat_stream(s, b, o, d, st.depth, stk, ctr, st.weight_threshold, st.max_steps,
                             [&](uint32_t cell, float near_t, float dt) {
      return tr.step(s, st, d, cell, near_t, dt);
    });
  tr.a.bytes += 64u * ctr.nodes + 32u * ctr.leaves + 32u * ctr.planes;
  store_out(out, idx, tr.a, tr.steps, tr.last, tr.t_end);
}

template <int DEG>
__global__ __launch_bounds__(kSplatBlock) void k_splat_relight(SceneView s, BvhView b, RaySource rs, Settings st,
                                                               bv_lighting lit, float* rgba_out) {
  extern __shared__ uint32_t lds_stack[];
  auto stk = reinterpret_cast<uint32_t (*)[kSplatBlock]>(lds_stack);
  const int64_t slot = (int64_t)blockIdx.x * kSplatBlock + threadIdx.x;
  if (slot >= ray_slots(rs)) return;
  V3 o, d;
  uint32_t start;
  int64_t idx;
  if (!fetch_ray(rs, slot, &o, &d, &start, &idx)) return;
  SplatCounters ctr{0u, 0u, 0u, 0u};
  float4 rgba = make_float4(0.f, 0.f, 0.f, 0.f);
  if (normalise(&d)) rgba = splat_relight<DEG>(s, b, st, lit, o, d, stk,
                                                reinterpret_cast<float*>(lds_stack + b.stack_rows * kSplatBlock), ctr);
  reinterpret_cast<float4*>(rgba_out)[idx] = rgba;
}

// Persistent serving kernel for splat requests over the scheduler's tile ring
// (ServeQueue): block 0 mirrors the host's counters, every other block (one
// wave) takes one 8x8 tile per ring entry and writes its pixels to the
// request's result buffer in pinned host memory.  At least 3 waves per SIMD
// (the relit walk's registers sit just above that bound otherwise).
template <int MODE, int DEG>
__global__ __launch_bounds__(kSplatBlock, 3) void k_serve_splat(ServeQueue q, uint32_t stack_rows) {
  // stack_rows x kSplatBlock stack entries (the deepest tree among the
  // server's scenes), then for relighting the per-light sums
  extern __shared__ uint32_t lds_stack[];
  auto stk = reinterpret_cast<uint32_t (*)[kSplatBlock]>(lds_stack);
  float* lsum = reinterpret_cast<float*>(lds_stack + stack_rows * kSplatBlock);
  __shared__ ServeReq req;
  __shared__ uint64_t entry_s;
  __shared__ uint32_t code_s;
  if (blockIdx.x == 0) {
    serve_mirror(q);
    return;
  }
  for (;;) {
    serve_claim(q, &entry_s, &code_s);
    const uint32_t code = code_s;
    const uint64_t entry = entry_s;
    if (code == kServeNoTile) break;
    serve_load_req(q, code >> 20, &req);
    const SceneView s = uniform_copy(req.s);
    const BvhView b = uniform_copy(req.bvh);
    const RaySource rs = uniform_copy(req.rs);
    const Settings st = uniform_copy(req.st);
    float* const out = uniform_copy(req.out);
    V3 o, d;
    uint32_t start;
    int64_t oi;
    const int64_t slot = (int64_t)(code & 0xFFFFFu) * 64 + threadIdx.x;
    if (fetch_ray(rs, slot, &o, &d, &start, &oi)) {
      float4 rgba = make_float4(0.f, 0.f, 0.f, 0.f);
      SplatCounters ctr{0u, 0u, 0u, 0u};
      if (normalise(&d)) {
        if (MODE == BV_EVAL_RELIGHT) {
          rgba = splat_relight<DEG>(s, b, st, req.lit, o, d, stk, lsum, ctr);
        } else {
          SplatTrace<BV_EVAL_SH, DEG> tr;
          tr.init(0u);
          splat_stream(s, b, o, d, st.depth, stk, ctr, st.weight_threshold, st.max_steps,
                                   [&](uint32_t cell, float near_t, float dt) {
            return tr.step(s, st, d, cell, near_t, dt);
          });
          rgba = make_float4(tr.a.r, tr.a.g, tr.a.b, 1.0f - tr.a.trans);
        }
      }
      reinterpret_cast<float4*>(out)[oi] = rgba;
    }
    serve_done(q, entry);
  }
}


}  // namespace bvk

// ---------------------------------------------------------------- host side

namespace bvh {

namespace {
std::mutex g_bvh_mutex;

}  // namespace

// Build (or rebuild after a geometry update) the sphere BVH of a power-cell scene.
int ensure_sphere_bvh(bv_scene* s, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_bvh_mutex);
  if (s->bvh_valid.load()) return BV_OK;
  const int64_t n = s->n;
  BV_REQUIRE(n > 0, BV_ERR_INVALID, "scene has no sites");
  if (!s->bvh_nodes) {
    const size_t node_bytes = (size_t)std::max<int64_t>(n - 1, 1) * sizeof(bvk::BvhNode);
    const size_t leaf_bytes = (size_t)n * sizeof(bvk::CellRec);
    BV_HIP(hipMalloc(&s->bvh_nodes, node_bytes));
    BV_HIP(hipMalloc(&s->bvh_leaf, leaf_bytes));
    BV_HIP(hipMalloc((void**)&s->bvh_cell, 4 * (size_t)n));
    s->device_bytes += (int64_t)(node_bytes + leaf_bytes + 4 * (size_t)n);
  }
  float4 *lo = nullptr, *hi = nullptr;
  BV_HIP(hipMalloc((void**)&lo, 16 * (size_t)n));
  if (hipMalloc((void**)&hi, 16 * (size_t)n) != hipSuccess) {
    (void)hipFree(lo);
    bvh::set_error("hipMalloc of BVH boxes failed");
    return BV_ERR_HIP;
  }
  const int bs = 256;
  const unsigned grid = (unsigned)((n + bs - 1) / bs);
  hipLaunchKernelGGL(bvk::k_sphere_boxes, dim3(grid), dim3(bs), 0, stream, n, s->pts, lo, hi);
  int rc = build_lbvh(n, s->pts, lo, hi, reinterpret_cast<bvk::BvhNode*>(s->bvh_nodes), s->bvh_cell, stream,
                      &s->bvh_depth);
  if (rc == BV_OK) {
    hipLaunchKernelGGL(bvk::k_leaves, dim3(grid), dim3(bs), 0, stream, n, s->bvh_cell, s->cells,
                       reinterpret_cast<bvk::CellRec*>(s->bvh_leaf));
    if (hipStreamSynchronize(stream) != hipSuccess) {
      bvh::set_error("BVH leaf gather failed");
      rc = BV_ERR_HIP;
    }
  }
  (void)hipFree(lo);
  (void)hipFree(hi);
  if (rc == BV_OK) s->bvh_valid.store(1);
  return rc;
}

int check_splat_scene(bv_scene* s, hipStream_t stream) {
  BV_REQUIRE(s != nullptr, BV_ERR_INVALID, "scene is null");
  BV_REQUIRE(s->bounded, BV_ERR_UNSUPPORTED, "PowerFoam splats require radii");
  return ensure_sphere_bvh(s, stream);
}

namespace {

bvk::BvhView bvh_view(const bv_scene* s) {
  return bvk::BvhView{reinterpret_cast<const bvk::BvhNode*>(s->bvh_nodes),
                      reinterpret_cast<const bvk::CellRec*>(s->bvh_leaf), s->bvh_cell,
                      bvk::stack_rows_for(s->bvh_depth)};
}

template <int MODE, int EVAL, int DEG, class RS>
void launch_splat_t(const bv_scene* s, const RS& rs, const bvk::Settings& st, const bvk::Outputs& out,
                    const bvk::SplatRecord& rec, const bvk::SplatJacRecord& jr, hipStream_t stream) {
  const int64_t slots = bvk::ray_slots(rs);
  const unsigned grid = (unsigned)((slots + bvk::kSplatBlock - 1) / bvk::kSplatBlock);
  const bvk::BvhView b = bvh_view(s);
  hipLaunchKernelGGL((bvk::k_splat<MODE, EVAL, DEG, RS>), dim3(grid), dim3(bvk::kSplatBlock),
                     bvk::stack_bytes(b.stack_rows), stream, s->view(), b, rs, st, out, rec, jr);
}

template <int MODE, int EVAL, class RS = bvk::RaySource>
void launch_splat_deg(const bv_scene* s, const RS& rs, const bvk::Settings& st, const bvk::Outputs& out,
                      const bvk::SplatRecord& rec, hipStream_t stream, const bvk::SplatJacRecord& jr = {}) {
  if constexpr (MODE != 0) {
    // recording reads no colour: one instantiation serves every degree
    launch_splat_t<MODE, EVAL, 0>(s, rs, st, out, rec, jr, stream);
  } else {
    switch (s->sh_degree < 3 ? s->sh_degree : 3) {
      case 0: launch_splat_t<MODE, EVAL, 0>(s, rs, st, out, rec, jr, stream); break;
      case 1: launch_splat_t<MODE, EVAL, 1>(s, rs, st, out, rec, jr, stream); break;
      case 2: launch_splat_t<MODE, EVAL, 2>(s, rs, st, out, rec, jr, stream); break;
      default: launch_splat_t<MODE, EVAL, 3>(s, rs, st, out, rec, jr, stream); break;
    }
  }
}

template <int DEG>
void launch_splat_relight_t(const bv_scene* s, const bvk::RaySource& rs, const bvk::Settings& st,
                            const bv_lighting& lit, float* out, hipStream_t stream) {
  const int64_t slots = bvk::ray_slots(rs);
  const unsigned grid = (unsigned)((slots + bvk::kSplatBlock - 1) / bvk::kSplatBlock);
  const bvk::BvhView b = bvh_view(s);
  hipLaunchKernelGGL((bvk::k_splat_relight<DEG>), dim3(grid), dim3(bvk::kSplatBlock),
                     bvk::stack_bytes(b.stack_rows) + bvk::kLightSumBytes, stream, s->view(), b, rs, st, lit, out);
}

int launch_splat_trace(bv_scene* s, const bvk::RaySource& rs, const bv_trace_settings* settings,
                       const bv_trace_outputs* outputs, hipStream_t stream) {
  BV_REQUIRE(settings && outputs, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(settings->max_steps > 0, BV_ERR_INVALID, "max_steps must be non-zero");
  BV_REQUIRE(settings->eval_mode <= BV_EVAL_SH, BV_ERR_INVALID, "eval_mode must be opacity, constant or SH");
  int rc = check_splat_scene(s, stream);
  if (rc != BV_OK) return rc;
  const bvk::Settings st = settings_of(settings);
  bvk::Outputs out{outputs->rgba, outputs->steps, outputs->last_point, outputs->t_end, outputs->depth_mode,
                   outputs->peak_weight, outputs->peak_point, outputs->stats};
  const bvk::SplatRecord rec{};
  if (bvk::ray_slots(rs) == 0) return BV_OK;
  switch (settings->eval_mode) {
    case BV_EVAL_OPACITY: launch_splat_deg<0, BV_EVAL_OPACITY>(s, rs, st, out, rec, stream); break;
    case BV_EVAL_CONSTANT: launch_splat_deg<0, BV_EVAL_CONSTANT>(s, rs, st, out, rec, stream); break;
    default: launch_splat_deg<0, BV_EVAL_SH>(s, rs, st, out, rec, stream); break;
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int launch_splat_relight(bv_scene* s, const bvk::RaySource& rs, const bv_trace_settings* settings,
                         const bv_lighting* lighting, float* rgba_out, hipStream_t stream) {
  BV_REQUIRE(settings && lighting, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(settings->max_steps > 0, BV_ERR_INVALID, "max_steps must be non-zero");
  BV_REQUIRE(lighting->count >= 1 && lighting->count <= BV_MAX_LIGHTS, BV_ERR_INVALID, "lighting needs 1..8 lights");
  int rc = check_splat_scene(s, stream);
  if (rc != BV_OK) return rc;
  BV_REQUIRE(s->material != nullptr, BV_ERR_INVALID, "relighting needs a scene with materials");
  if (bvk::ray_slots(rs) == 0) return BV_OK;
  BV_REQUIRE(rgba_out != nullptr, BV_ERR_INVALID, "rgba_out is null");
  const bvk::Settings st = settings_of(settings);
  switch (s->sh_degree < 3 ? s->sh_degree : 3) {
    case 0: launch_splat_relight_t<0>(s, rs, st, *lighting, rgba_out, stream); break;
    case 1: launch_splat_relight_t<1>(s, rs, st, *lighting, rgba_out, stream); break;
    case 2: launch_splat_relight_t<2>(s, rs, st, *lighting, rgba_out, stream); break;
    default: launch_splat_relight_t<3>(s, rs, st, *lighting, rgba_out, stream); break;
  }
  BV_HIP(hipGetLastError());
  return BV_OK;
}

}  // namespace

bvk::BvhView sphere_bvh_view(const bv_scene* s) { return bvh_view(s); }

int launch_splat_record_views(const bv_scene* s, const bvk::ViewRays& rs, uint32_t L, float depth, float max_dt,
                              const bv_splat_path_outputs& out, hipStream_t stream) {
  if (rs.count == 0) return BV_OK;
  bv_trace_settings ts{};
  ts.max_steps = L;
  ts.depth = depth;
  const bvk::SplatJacRecord jr{out.cells, out.next_cells, out.previous_cells, out.dts, out.mask,
                               reinterpret_cast<float4*>(out.dt_grad_previous),
                               reinterpret_cast<float4*>(out.dt_grad_current),
                               reinterpret_cast<float4*>(out.dt_grad_next), out.dt_reference_tangent, out.status, L,
                               max_dt};
  launch_splat_deg<2, BV_EVAL_OPACITY>(s, rs, settings_of(&ts), bvk::Outputs{}, bvk::SplatRecord{}, stream, jr);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

static const void* serve_splat_kernel_of(int mode, int sh_degree) {
  const int deg = sh_degree < 3 ? sh_degree : 3;
#define BV_SK(M)                                                                 \
  switch (deg) {                                                                 \
    case 0: return reinterpret_cast<const void*>(&bvk::k_serve_splat<M, 0>);     \
    case 1: return reinterpret_cast<const void*>(&bvk::k_serve_splat<M, 1>);     \
    case 2: return reinterpret_cast<const void*>(&bvk::k_serve_splat<M, 2>);     \
    default: return reinterpret_cast<const void*>(&bvk::k_serve_splat<M, 3>);    \
  }
  if (mode == BV_EVAL_RELIGHT) {
    BV_SK(BV_EVAL_RELIGHT)
  } else {
    BV_SK(BV_EVAL_SH)
  }
#undef BV_SK
}

static size_t serve_splat_lds(int mode, uint32_t stack_rows) {
  return bvk::stack_bytes(stack_rows) + (mode == BV_EVAL_RELIGHT ? bvk::kLightSumBytes : 0);
}

int serve_splat_blocks_per_cu(int mode, int sh_degree, uint32_t stack_rows) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, serve_splat_kernel_of(mode, sh_degree), bvk::kSplatBlock,
                                                   serve_splat_lds(mode, stack_rows)) != hipSuccess ||
      b < 1)
    b = 1;
  return b;
}

int launch_serve_splat(int mode, int sh_degree, const bvk::ServeQueue& q, uint32_t blocks, uint32_t stack_rows,
                       hipStream_t stream) {
  bvk::ServeQueue qa = q;
  uint32_t rows = stack_rows;
  void* args[] = {&qa, &rows};
  BV_HIP(hipLaunchKernel(serve_splat_kernel_of(mode, sh_degree), dim3(blocks), dim3(bvk::kSplatBlock), args,
                         serve_splat_lds(mode, stack_rows), stream));
  return BV_OK;
}

}  // namespace bvh

extern "C" {

int bv_splat_trace_rays(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                        const bv_trace_outputs* outputs, void* stream) {
  BV_REQUIRE(rays != nullptr, BV_ERR_INVALID, "rays is null");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(rays->count == 0 || (rays->origins && rays->directions), BV_ERR_INVALID, "null ray arrays");
  bvk::RaySource rs = bvh::ray_source(rays);
  return bvh::launch_splat_trace(scene, rs, settings, outputs, (hipStream_t)stream);
}

int bv_splat_trace_camera(bv_scene* scene, const bv_camera* camera, uint32_t width, uint32_t height,
                          const bv_trace_settings* settings, const bv_trace_outputs* outputs, void* stream) {
  BV_REQUIRE(camera != nullptr, BV_ERR_INVALID, "camera is null");
  BV_REQUIRE(width > 0 && height > 0, BV_ERR_INVALID, "resolution must be non-zero");
  bvk::RaySource rs = bvh::camera_source(camera, width, height, 0u);
  return bvh::launch_splat_trace(scene, rs, settings, outputs, (hipStream_t)stream);
}

int bv_splat_record(bv_scene* scene, const bv_rays* rays, uint32_t max_steps, float depth, uint32_t* cells,
                    float* nears, float* dts, uint32_t* count, void* stream) {
  BV_REQUIRE(rays != nullptr, BV_ERR_INVALID, "rays is null");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(rays->count == 0 || (rays->origins && rays->directions), BV_ERR_INVALID, "null ray arrays");
  BV_REQUIRE(max_steps > 0, BV_ERR_INVALID, "max_steps must be non-zero");
  BV_REQUIRE(rays->count == 0 || (cells && nears && dts && count), BV_ERR_INVALID, "null record buffers");
  hipStream_t st = (hipStream_t)stream;
  int rc = bvh::check_splat_scene(scene, st);
  if (rc != BV_OK) return rc;
  if (rays->count == 0) return BV_OK;
  bvk::RaySource rs = bvh::ray_source(rays);
  bv_trace_settings ts{};
  ts.max_steps = max_steps;
  ts.depth = depth;
  const bvk::Settings se = bvh::settings_of(&ts);
  const bvk::Outputs out{};
  const bvk::SplatRecord rec{cells, nears, dts, count, max_steps};
  bvh::launch_splat_deg<1, BV_EVAL_OPACITY>(scene, rs, se, out, rec, st);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int bv_splat_record_jacobians(bv_scene* scene, const bv_rays* rays, const bv_record_settings* settings,
                              const bv_splat_path_outputs* out, void* stream) {
  BV_REQUIRE(rays != nullptr && settings != nullptr && out != nullptr, BV_ERR_INVALID, "null argument");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(rays->count == 0 || (rays->origins && rays->directions), BV_ERR_INVALID, "null ray arrays");
  BV_REQUIRE(settings->max_steps > 0, BV_ERR_INVALID, "max_steps must be non-zero");
  BV_REQUIRE(rays->count == 0 || (out->cells && out->next_cells && out->previous_cells && out->dts &&
                                  out->dt_grad_previous && out->dt_grad_current && out->dt_grad_next &&
                                  out->dt_reference_tangent && out->status),
             BV_ERR_INVALID, "null record buffers");
  hipStream_t st = (hipStream_t)stream;
  int rc = bvh::check_splat_scene(scene, st);
  if (rc != BV_OK) return rc;
  if (rays->count == 0) return BV_OK;
  bvk::RaySource rs = bvh::ray_source(rays);
  bv_trace_settings ts{};
  ts.max_steps = settings->max_steps;
  ts.depth = settings->depth;
  const bvk::Settings se = bvh::settings_of(&ts);
  const bvk::SplatJacRecord jr{out->cells, out->next_cells, out->previous_cells, out->dts, out->mask,
                               reinterpret_cast<float4*>(out->dt_grad_previous),
                               reinterpret_cast<float4*>(out->dt_grad_current),
                               reinterpret_cast<float4*>(out->dt_grad_next), out->dt_reference_tangent, out->status,
                               settings->max_steps, settings->max_path_dt};
  bvh::launch_splat_deg<2, BV_EVAL_OPACITY>(scene, rs, se, bvk::Outputs{}, bvk::SplatRecord{}, st, jr);
  BV_HIP(hipGetLastError());
  return BV_OK;
}

int bv_splat_relight_camera(bv_scene* scene, const bv_camera* camera, uint32_t width, uint32_t height,
                            const bv_trace_settings* settings, const bv_lighting* lighting, float* rgba_out,
                            void* stream) {
  BV_REQUIRE(camera != nullptr, BV_ERR_INVALID, "camera is null");
  BV_REQUIRE(width > 0 && height > 0, BV_ERR_INVALID, "resolution must be non-zero");
  bvk::RaySource rs = bvh::camera_source(camera, width, height, 0u);
  return bvh::launch_splat_relight(scene, rs, settings, lighting, rgba_out, (hipStream_t)stream);
}

int bv_splat_relight_rays(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                          const bv_lighting* lighting, float* rgba_out, void* stream) {
  BV_REQUIRE(rays != nullptr, BV_ERR_INVALID, "rays is null");
  BV_REQUIRE(rays->count >= 0, BV_ERR_INVALID, "negative ray count");
  BV_REQUIRE(rays->count == 0 || (rays->origins && rays->directions), BV_ERR_INVALID, "null ray arrays");
  bvk::RaySource rs = bvh::ray_source(rays);
  return bvh::launch_splat_relight(scene, rs, settings, lighting, rgba_out, (hipStream_t)stream);
}

}  // extern "C"
