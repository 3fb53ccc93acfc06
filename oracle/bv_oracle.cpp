// CPU oracle for the RadFoam / PowerFoam cell walk.  TEST INFRASTRUCTURE ONLY:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker.  Nothing in the product links it.
//
// A from-scratch restatement of blade-volume/src/trace.rs (the reference's CPU
// tracer, which mirrors shaders/radfoam_trace.wgsl step for step):
//   find_exit_face            trace.rs:170-205
//   support_interval          trace.rs:209-238
//   SH basis + evaluation     trace.rs:242-402 (term order of trace.rs, not sh_eval.wgsl)
//   trace_one_ray             trace.rs:1680-1820
//   record_paths              shaders/radfoam_record_paths.wgsl:1220-1355
//                             (GPU recorder: status word, max_path_dt clamp)
//   containing_cell           lib.rs:526-541
//   camera ray                shaders/radfoam.wgsl:45-52 + common.wgsl:19-23 (qrot)
//   recorder ray              shaders/radfoam_record_paths.wgsl:133-153 (its own qrot)
//   point-light shading       relight.rs:203-324 (PointLight::sample, point_light_specular,
//                             shade_point_light); trace_relight composes it with the walk
//   adjoint                   gradient of trace_one_ray (SH mode) w.r.t. density,
//                             SH coefficients and site positions, following the
//                             trainer graph blade-volume-train/src/diff_render.rs:1306-1572
//                             (dt from bisector planes, T/alpha/weight) and the
//                             closed-form face Jacobian of trace.rs:1077-1097.
//
// Every function is templated on the scalar type: float is the parity oracle,
// double is the twin used for central finite differences.  Compile with
// -ffp-contract=off and no fast-math: the reference is Rust, which never
// contracts a*b+c into an FMA, and glam's Vec3 ops evaluate left to right.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> struct V3 { T x, y, z; };
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// glam Vec3::dot: (x*x') + (y*y') + (z*z'), evaluated left to right.
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float exp_t(float x) { return std::exp(x); }   // libm expf, as Rust f32::exp on Linux
inline double exp_t(double x) { return std::exp(x); }
inline float sqrt_t(float x) { return std::sqrt(x); }
inline double sqrt_t(double x) { return std::sqrt(x); }

template <class T> struct Scene {
  int64_t n;
  const T* xyz;        // [n][3]
  const T* radius;     // [n] or null (null = plain RadFoam, unbounded cells)
  const T* density;    // [n]
  const T* sh;         // [n][3*K], interleaved RGB per component (trace.rs:271-272)
  int32_t sh_degree;
  const uint32_t* offsets;    // [n+1]
  const uint32_t* neighbors;  // [offsets[n]]
};

template <class T> inline V3<T> site(const Scene<T>& s, uint32_t i) {
  return v3<T>(s.xyz[3 * (size_t)i], s.xyz[3 * (size_t)i + 1], s.xyz[3 * (size_t)i + 2]);
}
template <class T> inline T radius_of(const Scene<T>& s, uint32_t i) { return s.radius ? s.radius[i] : T(0); }

// trace.rs:242-261.  Stored as f32 literals; the double twin promotes them so
// it differentiates the very same function.
static const float kShC[16] = {
    0.2820948f,  -0.48860252f, 0.48860252f, -0.48860252f, 1.0925485f, -1.0925485f,
    0.31539157f, -1.0925485f,  0.54627424f, -0.5900436f,  2.8906114f, -0.4570458f,
    0.37317634f, -0.4570458f,  1.4453057f,  -0.5900436f};

inline int sh_components(int deg) { int d = (deg < 3 ? deg : 3) + 1; return d * d; }

// Real SH basis value of component k in direction (x,y,z), without the constant:
// the factor trace.rs multiplies the (constant * coefficient) product by.
// Returns the list in component order; trace.rs:287-392.
template <class T> inline void sh_factors(V3<T> d, T f[16]) {
  const T x = d.x, y = d.y, z = d.z;
  const T xx = x * x, yy = y * y, zz = z * z;
  f[0] = T(1);
  f[1] = y; f[2] = z; f[3] = x;
  f[4] = 0; f[5] = 0; f[6] = T(3) * zz - T(1); f[7] = 0; f[8] = xx - yy;  // 4,5,7: two-factor products, see eval
  f[9] = y * (T(3) * xx - yy);
  f[10] = x * y * z;
  f[11] = y * (T(5) * zz - T(1));
  f[12] = z * (T(5) * zz - T(3));
  f[13] = x * (T(5) * zz - T(1));
  f[14] = z * (xx - yy);
  f[15] = x * (xx - T(3) * yy);
}

// eval_rgb_sh_unclamped, trace.rs:264-392: color starts at zero, each term is
// (C_k * coef) * factor (degree-2 xy/yz/xz terms are ((C_k*coef)*a)*b), then 0.5 + color.
template <class T> V3<T> eval_sh_unclamped(const Scene<T>& s, uint32_t cell, V3<T> dir) {
  const int deg = s.sh_degree < 3 ? s.sh_degree : 3;
  const int comps = sh_components(deg);
  const T* c = s.sh + (size_t)cell * (size_t)(3 * comps);
  T f[16];
  sh_factors(dir, f);
  T r = 0, g = 0, b = 0;
  r += T(kShC[0]) * c[0]; g += T(kShC[0]) * c[1]; b += T(kShC[0]) * c[2];
  if (deg >= 1) {
    for (int k = 1; k <= 3; ++k) {
      r += T(kShC[k]) * c[3 * k] * f[k]; g += T(kShC[k]) * c[3 * k + 1] * f[k]; b += T(kShC[k]) * c[3 * k + 2] * f[k];
    }
  }
  if (deg >= 2) {
    const T x = dir.x, y = dir.y, z = dir.z;
    // k=4: x*y, k=5: y*z, k=7: x*z evaluated as ((C*coef)*a)*b
    r += T(kShC[4]) * c[12] * x * y; g += T(kShC[4]) * c[13] * x * y; b += T(kShC[4]) * c[14] * x * y;
    r += T(kShC[5]) * c[15] * y * z; g += T(kShC[5]) * c[16] * y * z; b += T(kShC[5]) * c[17] * y * z;
    r += T(kShC[6]) * c[18] * f[6]; g += T(kShC[6]) * c[19] * f[6]; b += T(kShC[6]) * c[20] * f[6];
    r += T(kShC[7]) * c[21] * x * z; g += T(kShC[7]) * c[22] * x * z; b += T(kShC[7]) * c[23] * x * z;
    r += T(kShC[8]) * c[24] * f[8]; g += T(kShC[8]) * c[25] * f[8]; b += T(kShC[8]) * c[26] * f[8];
  }
  if (deg >= 3) {
    for (int k = 9; k <= 15; ++k) {
      r += T(kShC[k]) * c[3 * k] * f[k]; g += T(kShC[k]) * c[3 * k + 1] * f[k]; b += T(kShC[k]) * c[3 * k + 2] * f[k];
    }
  }
  return v3<T>(T(0.5) + r, T(0.5) + g, T(0.5) + b);
}

// Basis value multiplying coefficient k (including the constant), for the adjoint:
// d color / d coef_k = C_k * factor_k.
template <class T> inline void sh_basis(V3<T> d, int comps, T out[16]) {
  T f[16];
  sh_factors(d, f);
  f[4] = d.x * d.y; f[5] = d.y * d.z; f[7] = d.x * d.z;
  for (int k = 0; k < comps; ++k) out[k] = T(kShC[k]) * f[k];
}

template <class T> struct Settings {
  T weight_threshold;
  uint32_t max_steps;
  T depth;
  uint32_t eval_mode;  // 0 = opacity, 1 = constant rgb, 2 = SH (trace.rs:38-48)
  T const_rgb[3];
};

template <class T> struct TraceOut {
  T rgba[4];
  uint32_t steps;
  uint32_t last_point;
  T t_end;
  T depth_mode;
  T peak_weight;
  uint32_t peak_point;  // 0xFFFFFFFF = none
};

// find_exit_face, trace.rs:170-205 (both the weighted and unweighted branch).
template <class T>
inline void find_exit_face(const Scene<T>& s, uint32_t cell, V3<T> cur, T cur_r, V3<T> ro, V3<T> rd, T t0,
                           T depth, T* best_t1, int64_t* face) {
  T best = depth;
  int64_t nf = -1;
  const T r_i_sq = cur_r * cur_r;
  const uint32_t b = s.offsets[cell], e = s.offsets[cell + 1];
  for (uint32_t j = b; j < e; ++j) {
    const uint32_t nb = s.neighbors[j];
    const V3<T> off = site(s, nb) - cur;
    T shift;
    if (s.radius) {
      const T r_j = s.radius[nb];
      T dsq = dot(off, off);
      dsq = dsq > T(1e-20f) ? dsq : T(1e-20f);  // .max(1e-20)
      shift = T(0.5) + T(0.5) * (r_i_sq - r_j * r_j) / dsq;
    } else {
      shift = T(0.5);
    }
    const V3<T> fo = cur + shift * off;
    const T dp = dot(off, rd);
    if (dp > T(0)) {
      const T t = dot(fo - ro, off) / dp;
      if (std::isfinite(t) && t > t0 && t < best) {
        best = t;
        nf = (int64_t)(j - b);
      }
    }
  }
  *best_t1 = best;
  *face = nf;
}

// support_interval, trace.rs:209-238.
template <class T>
inline bool support_interval(bool bounded, V3<T> ro, V3<T> rd, V3<T> c, T radius, T t0, T t1, T* s0, T* s1) {
  if (!bounded) {
    *s0 = t0; *s1 = t1;
    return t1 > t0;
  }
  const V3<T> oc = ro - c;
  const T a = dot(rd, rd);
  if (!std::isfinite(a) || a <= T(0)) return false;
  const T b = dot(oc, rd);
  const T cc = dot(oc, oc) - radius * radius;
  const T disc = b * b - a * cc;
  if (disc <= T(0)) return false;
  const T root = sqrt_t(disc);
  const T near_t = (-b - root) / a;
  const T far_t = (-b + root) / a;
  const T cs = t0 > near_t ? t0 : near_t;  // t0.max(near)
  const T ce = t1 < far_t ? t1 : far_t;    // t1.min(far)
  *s0 = cs; *s1 = ce;
  return ce > cs;
}

// Normalise a direction the way trace.rs:1704-1706 does (divide by length).
template <class T> inline bool normalise(V3<T>* d) {
  const T len = sqrt_t(dot(*d, *d));
  if (!(len > T(0)) || !std::isfinite(len)) return false;
  *d = *d / len;
  return true;
}

// trace_one_ray, trace.rs:1680-1820.  For object scenes (keep_length) the walk
// runs along dir_in as given, so t is the caller's parameter, while SH uses the
// normalised direction; win integrates only from t_lo on (integration_begin =
// max(support start, integration_start), radfoam_trace.wgsl:159-160).
template <class T>
TraceOut<T> trace_one_ray(const Scene<T>& s, V3<T> ro, V3<T> dir_in, uint32_t start, const Settings<T>& st,
                          bool keep_length = false, bool win = false, T t_lo = T(0), uint32_t* path = nullptr,
                          uint32_t path_len = 0) {
  TraceOut<T> r{};
  r.last_point = start;
  r.peak_point = 0xFFFFFFFFu;
  V3<T> sdir = dir_in;
  if (!normalise(&sdir)) return r;
  const V3<T> dir = keep_length ? dir_in : sdir;
  T t0 = 0, trans = 1, ar = 0, ag = 0, ab = 0, depth_mode = 0, peak = 0;
  uint32_t peak_point = 0xFFFFFFFFu;
  uint32_t cur = start;
  V3<T> cur_pos = site(s, cur);
  T cur_r = radius_of(s, cur);
  const bool bounded = s.radius != nullptr;
  uint32_t steps = 0;
  while (steps < st.max_steps) {
    steps += 1;
    // the visited-cell sequence: the current cell of every iteration
    if (path != nullptr && steps - 1 < path_len) path[steps - 1] = cur;
    if (trans <= st.weight_threshold || t0 > st.depth) break;
    T best_t1;
    int64_t face;
    find_exit_face(s, cur, cur_pos, cur_r, ro, dir, t0, st.depth, &best_t1, &face);
    T seg0, seg1;
    bool seg = support_interval(bounded, ro, dir, cur_pos, cur_r, t0, best_t1, &seg0, &seg1);
    if (seg && win) {
      seg0 = seg0 > t_lo ? seg0 : t_lo;
      seg = seg1 > seg0;
    }
    if (seg) {
      const T sd = s.density[cur];
      if (sd > T(1e-6f)) {
        const T dt = seg1 - seg0;
        const T alpha = T(1) - exp_t(-sd * dt);
        const T w = trans * alpha;
        if (st.eval_mode == 1) {
          ar += w * st.const_rgb[0]; ag += w * st.const_rgb[1]; ab += w * st.const_rgb[2];
        } else if (st.eval_mode == 2) {
          V3<T> c = eval_sh_unclamped(s, cur, sdir);
          c.x = c.x > T(0) ? c.x : T(0); c.y = c.y > T(0) ? c.y : T(0); c.z = c.z > T(0) ? c.z : T(0);
          ar += w * c.x; ag += w * c.y; ab += w * c.z;
        }
        if (w > peak) {
          peak = w;
          depth_mode = T(0.5) * (seg0 + seg1);
          peak_point = cur;
        }
        trans *= T(1) - alpha;
      }
    }
    if (face < 0) {
      t0 = best_t1;
      break;
    }
    const uint32_t nb = s.neighbors[s.offsets[cur] + (uint32_t)face];
    t0 = t0 > best_t1 ? t0 : best_t1;  // t0.max(best_t1)
    cur = nb;
    cur_pos = site(s, nb);
    cur_r = radius_of(s, nb);
  }
  r.rgba[0] = ar; r.rgba[1] = ag; r.rgba[2] = ab; r.rgba[3] = T(1) - trans;
  r.steps = steps;
  r.last_point = cur;
  r.t_end = t0;
  r.depth_mode = depth_mode;
  r.peak_weight = peak;
  r.peak_point = peak_point;
  return r;
}

// GPU path recorder, radfoam_record_paths.wgsl:1220-1355 with the unweighted
// interval_differential of :291-305.  Rows are [P, L]; trailing slots of a row
// are not written.  status = count | (truncated << 31).
template <class T>
uint32_t record_path(const Scene<T>& s, V3<T> ro, V3<T> dir_in, uint32_t start, uint32_t L, T depth, T max_dt,
                     uint32_t* cells, uint32_t* next_cells, T* dts) {
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return 0;
  T t0 = 0;
  uint32_t cur = start;
  V3<T> cur_pos = site(s, cur);
  T cur_r = radius_of(s, cur);
  uint32_t out = 0;
  bool truncated = false;
  for (uint32_t step = 0; step < L; ++step) {
    T t1;
    int64_t face;
    find_exit_face(s, cur, cur_pos, cur_r, ro, dir, t0, depth, &t1, &face);
    if (t1 <= t0) break;
    const uint32_t nxt = face >= 0 ? s.neighbors[s.offsets[cur] + (uint32_t)face] : cur;
    // Unweighted recorder: the interval is the whole cell interval.
    T dt = t1 - t0;
    dt = dt < max_dt ? dt : max_dt;  // min(t1 - t0, max_path_dt)
    cells[out] = cur;
    next_cells[out] = nxt;
    dts[out] = dt;
    out += 1;
    if (face < 0) break;
    if (step + 1 == L) truncated = true;
    t0 = t1;
    cur = nxt;
    cur_pos = site(s, nxt);
    cur_r = radius_of(s, nxt);
  }
  return out | (truncated ? 0x80000000u : 0u);
}

// Closed-form derivative of one bisector-plane crossing t with respect to the
// two sites, trace.rs:1077-1097 (radius terms dropped: plain RadFoam).
template <class T>
inline void face_jac(V3<T> ro, V3<T> rd, V3<T> cur, V3<T> adj, T t, V3<T>* d_cur, V3<T>* d_adj) {
  const V3<T> nrm = adj - cur;
  const T den = dot(rd, nrm);
  if (!std::isfinite(den) || std::fabs(den) <= T(1e-20f)) {
    *d_cur = v3<T>(0, 0, 0);
    *d_adj = v3<T>(0, 0, 0);
    return;
  }
  *d_cur = (ro - cur + t * rd) / den;
  *d_adj = (adj - ro - t * rd) / den;
}

// Adjoint of trace_one_ray in SH mode for plain RadFoam.  Accumulates
// d_density[N], d_sh[N*3K], d_xyz[N*3] given dL/d(rgba).  One forward pass
// yields C and T_final; one replay then applies, per integrated segment k,
//   dL/dtau_k = sum_c g_c (T_{k+1} c_k - (C - C_{<=k})) + g_a T_final,
// tau_k = sigma_k dt_k, dt_k = t_exit - t_entry of the bisector-plane crossings.
// Gradient accumulation: plain for the (serial) parity oracle, an OpenMP
// atomic for the multi-threaded timing baseline (x - y == x + (-y) exactly).
template <bool ATOMIC, class T> inline void add_to(T* p, T v) {
  if constexpr (ATOMIC) {
#pragma omp atomic
    *p += v;
  } else {
    *p += v;
  }
}

template <class T, bool ATOMIC = false>
void adjoint_one_ray(const Scene<T>& s, V3<T> ro, V3<T> dir_in, uint32_t start, const Settings<T>& st,
                     const T g[4], T* d_density, T* d_sh, T* d_xyz) {
  const TraceOut<T> fwd = trace_one_ray(s, ro, dir_in, start, st);
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return;
  const int comps = sh_components(s.sh_degree);
  T basis[16];
  sh_basis(dir, comps, basis);
  const T Cr = fwd.rgba[0], Cg = fwd.rgba[1], Cb = fwd.rgba[2];
  const T Tfin = T(1) - fwd.rgba[3];
  T t0 = 0, trans = 1, pr = 0, pg = 0, pb = 0;
  uint32_t cur = start, prev = start;
  V3<T> cur_pos = site(s, cur);
  uint32_t steps = 0;
  while (steps < st.max_steps) {
    steps += 1;
    if (trans <= st.weight_threshold || t0 > st.depth) break;
    T t1;
    int64_t face;
    find_exit_face(s, cur, cur_pos, T(0), ro, dir, t0, st.depth, &t1, &face);
    const uint32_t nxt = face >= 0 ? s.neighbors[s.offsets[cur] + (uint32_t)face] : cur;
    if (t1 > t0) {
      const T sd = s.density[cur];
      if (sd > T(1e-6f)) {
        const T dt = t1 - t0;
        const T alpha = T(1) - exp_t(-sd * dt);
        const T w = trans * alpha;
        const V3<T> cu = eval_sh_unclamped(s, cur, dir);
        const T cr = cu.x > T(0) ? cu.x : T(0), cg = cu.y > T(0) ? cu.y : T(0), cb = cu.z > T(0) ? cu.z : T(0);
        pr += w * cr; pg += w * cg; pb += w * cb;
        const T tn = trans * (T(1) - alpha);
        const T gtau = g[0] * (tn * cr - (Cr - pr)) + g[1] * (tn * cg - (Cg - pg)) + g[2] * (tn * cb - (Cb - pb)) +
                       g[3] * Tfin;
        add_to<ATOMIC>(&d_density[cur], gtau * dt);
        const T gdt = gtau * sd;
        // SH: d rgb_c / d coef(k, c) = basis_k where the channel is unclamped.
        T* dsh = d_sh + (size_t)cur * (size_t)(3 * comps);
        const T gc[3] = {cu.x > T(0) ? g[0] * w : T(0), cu.y > T(0) ? g[1] * w : T(0), cu.z > T(0) ? g[2] * w : T(0)};
        for (int k = 0; k < comps; ++k)
          for (int c = 0; c < 3; ++c) add_to<ATOMIC>(&dsh[3 * k + c], gc[c] * basis[k]);
        // Positions through dt = t1 - t0.
        if (face >= 0) {
          V3<T> dc, dn;
          face_jac(ro, dir, cur_pos, site(s, nxt), t1, &dc, &dn);
          T* xc = d_xyz + 3 * (size_t)cur;
          T* xn = d_xyz + 3 * (size_t)nxt;
          add_to<ATOMIC>(xc, gdt * dc.x); add_to<ATOMIC>(xc + 1, gdt * dc.y); add_to<ATOMIC>(xc + 2, gdt * dc.z);
          add_to<ATOMIC>(xn, gdt * dn.x); add_to<ATOMIC>(xn + 1, gdt * dn.y); add_to<ATOMIC>(xn + 2, gdt * dn.z);
        }
        if (prev != cur) {
          V3<T> dp, dc;
          face_jac(ro, dir, site(s, prev), cur_pos, t0, &dp, &dc);
          T* xp = d_xyz + 3 * (size_t)prev;
          T* xc = d_xyz + 3 * (size_t)cur;
          add_to<ATOMIC>(xp, -(gdt * dp.x)); add_to<ATOMIC>(xp + 1, -(gdt * dp.y)); add_to<ATOMIC>(xp + 2, -(gdt * dp.z));
          add_to<ATOMIC>(xc, -(gdt * dc.x)); add_to<ATOMIC>(xc + 1, -(gdt * dc.y)); add_to<ATOMIC>(xc + 2, -(gdt * dc.z));
        }
        trans = tn;
      }
    }
    if (face < 0) break;
    t0 = t0 > t1 ? t0 : t1;
    prev = cur;
    cur = nxt;
    cur_pos = site(s, nxt);
  }
}

// ---------------------------------------------------------------- PowerFoam splats
// One support interval of a PowerFoam splat: the part of the cell's support
// sphere that survives every radical plane of its (Cech) adjacency row.
template <class T> struct SplatInterval {
  T near_t, dt;
  uint32_t cell;
};

// sphere_intersection_jacobians (trace.rs:1101-1124) without the Jacobians:
// oc = o - c, b = oc.d, disc = b*b - (|oc|^2 - r*r), roots -b -+ sqrt(disc).
template <class T> inline bool sphere_hit(V3<T> ro, V3<T> rd, V3<T> c, T r, T* near_t, T* far_t) {
  const V3<T> oc = ro - c;
  const T b = dot(oc, rd);
  const T disc = b * b - (dot(oc, oc) - r * r);
  if (disc <= T(0) || !std::isfinite(disc)) return false;
  const T root = sqrt_t(disc);
  *near_t = -b - root;
  *far_t = -b + root;
  return true;
}

// powerfoam_splat_interval (trace.rs:1331-1389) followed by the bounded branch of
// path_interval_jacobian (trace.rs:1199-1318, no oriented surface): camera
// exclusion at 4 r, sphere hit inside (0, depth), then the entry/exit faces of
// every neighbour plane (denominators beyond +-1e-20; a parallel plane with
// the ray on its far side empties the interval), clipped to the sphere chord.
template <class T>
bool splat_interval(const Scene<T>& s, V3<T> ro, V3<T> rd, T depth, uint32_t cell, SplatInterval<T>* out) {
  const V3<T> cur = site(s, cell);
  const T cr = radius_of(s, cell);
  const V3<T> cam = cur - ro;
  const T excl = T(4) * cr;
  if (dot(cam, cam) < excl * excl) return false;
  T sn, sf;
  if (!sphere_hit(ro, rd, cur, cr, &sn, &sf)) return false;
  if (sf <= T(0) || sn >= depth) return false;
  T fn = T(0), ff = depth;
  for (uint32_t e = s.offsets[cell]; e < s.offsets[cell + 1]; ++e) {
    const uint32_t j = s.neighbors[e];
    const T rj = radius_of(s, j);
    const V3<T> nrm = site(s, j) - cur;
    T dsq = dot(nrm, nrm);
    dsq = dsq > T(1e-20f) ? dsq : T(1e-20f);
    const T shift = T(0.5) + T(0.5) * (cr * cr - rj * rj) / dsq;
    const V3<T> fo = cur + shift * nrm;
    const T num = dot(fo - ro, nrm);
    const T den = dot(rd, nrm);
    if (den > T(1e-20f)) {
      const T t = num / den;
      if (t < ff) ff = t;
    } else if (den < T(-1e-20f)) {
      const T t = num / den;
      if (t > fn) fn = t;
    } else if (num < T(0)) {
      return false;
    }
  }
  T start = fn, end = ff;
  if (sn > start) start = sn;
  if (sf < end) end = sf;
  if (end <= start || !std::isfinite(start) || !std::isfinite(end)) return false;
  out->near_t = start;
  out->dt = end - start;
  out->cell = cell;
  return true;
}

// sorted_powerfoam_splat_intervals (trace.rs:1391-1411): every cell's interval,
// ordered by (near, cell) and truncated to max_steps.  rd is already unit.
template <class T>
void sorted_splats(const Scene<T>& s, V3<T> ro, V3<T> rd, T depth, uint32_t max_steps,
                   std::vector<SplatInterval<T>>* out) {
  out->clear();
  for (int64_t c = 0; c < s.n; ++c) {
    SplatInterval<T> iv;
    if (splat_interval(s, ro, rd, depth, (uint32_t)c, &iv)) out->push_back(iv);
  }
  std::sort(out->begin(), out->end(), [](const SplatInterval<T>& a, const SplatInterval<T>& b) {
    return a.near_t < b.near_t || (a.near_t == b.near_t && a.cell < b.cell);
  });
  if (out->size() > max_steps) out->resize(max_steps);
}

// One interval with the exact local derivatives of its length (the PowerFoam
// recorder's Jacobian streams): path_interval_jacobian (trace.rs:1199-1318,
// bounded branch, no oriented surface) on the faces powerfoam_splat_interval
// (:1331-1389) picked, as the GPU recorder forms them (interval_differential,
// radfoam_record_paths.wgsl:291-430) including its clamp to max_path_dt with
// zero derivatives on the saturated branch, plus the reference tangent written
// beside them (:1064-1087).
template <class T> struct SplatJac {
  T near_t, dt;
  uint32_t cell, prev, next;
  T jp[4], jc[4], jn[4];
  T tangent;
};

// face_intersection_jacobians, trace.rs:1077-1097.
template <class T>
void face_jac4(V3<T> ro, V3<T> rd, V3<T> cur, T rc, V3<T> adj, T ra, T t, T jc[4], T ja[4]) {
  const V3<T> nrm = adj - cur;
  const T den = dot(rd, nrm);
  for (int q = 0; q < 4; ++q) jc[q] = ja[q] = T(0);
  if (!std::isfinite(den) || std::fabs(den) <= T(1e-20f)) return;
  const V3<T> a = (ro - cur) + t * rd;
  const V3<T> b = (adj - ro) - t * rd;
  jc[0] = a.x / den; jc[1] = a.y / den; jc[2] = a.z / den; jc[3] = rc / den;
  ja[0] = b.x / den; ja[1] = b.y / den; ja[2] = b.z / den; ja[3] = -ra / den;
}

// sphere_intersection_jacobians, trace.rs:1101-1124 (d unit).
template <class T>
bool sphere_jac4(V3<T> ro, V3<T> rd, V3<T> c, T r, T* near_t, T* far_t, T nj[4], T fj[4]) {
  const V3<T> oc = ro - c;
  const T b = dot(oc, rd);
  const T disc = b * b - (dot(oc, oc) - r * r);
  if (disc <= T(0) || !std::isfinite(disc)) return false;
  const T root = sqrt_t(disc);
  const V3<T> perp = oc - b * rd;
  const V3<T> rdc = perp / root;
  const T rdr = r / root;
  *near_t = -b - root;
  *far_t = -b + root;
  nj[0] = rd.x - rdc.x; nj[1] = rd.y - rdc.y; nj[2] = rd.z - rdc.z; nj[3] = -rdr;
  fj[0] = rd.x + rdc.x; fj[1] = rd.y + rdc.y; fj[2] = rd.z + rdc.z; fj[3] = rdr;
  return true;
}

template <class T> inline T dot4(const T a[4], V3<T> g, T w) { return ((a[0] * g.x + a[1] * g.y) + a[2] * g.z) + a[3] * w; }

template <class T>
bool splat_interval_jac(const Scene<T>& s, V3<T> ro, V3<T> rd, T depth, uint32_t cell, T max_dt, SplatJac<T>* out) {
  const V3<T> cur = site(s, cell);
  const T cr = radius_of(s, cell);
  const V3<T> cam = cur - ro;
  const T excl = T(4) * cr;
  if (dot(cam, cam) < excl * excl) return false;
  T sn, sf, nj[4], fj[4];
  if (!sphere_jac4(ro, rd, cur, cr, &sn, &sf, nj, fj)) return false;
  if (sf <= T(0) || sn >= depth) return false;
  T fn = T(0), ff = depth;
  uint32_t prev = cell, next = cell;
  for (uint32_t e = s.offsets[cell]; e < s.offsets[cell + 1]; ++e) {
    const uint32_t j = s.neighbors[e];
    const T rj = radius_of(s, j);
    const V3<T> nrm = site(s, j) - cur;
    T dsq = dot(nrm, nrm);
    dsq = dsq > T(1e-20f) ? dsq : T(1e-20f);
    const T shift = T(0.5) + T(0.5) * (cr * cr - rj * rj) / dsq;
    const V3<T> fo = cur + shift * nrm;
    const T num = dot(fo - ro, nrm);
    const T den = dot(rd, nrm);
    if (den > T(1e-20f)) {
      const T t = num / den;
      if (t < ff) { ff = t; next = j; }
    } else if (den < T(-1e-20f)) {
      const T t = num / den;
      if (t > fn) { fn = t; prev = j; }
    } else if (num < T(0)) {
      return false;
    }
  }
  T start = fn, end = ff;
  const bool start_sphere = sn > start;
  if (start_sphere) start = sn;
  const bool end_sphere = sf < end;
  if (end_sphere) end = sf;
  if (end <= start || !std::isfinite(start) || !std::isfinite(end)) return false;
  SplatJac<T> r{};
  r.near_t = start;
  r.cell = cell;
  r.prev = prev;
  r.next = next;
  if (start_sphere) {
    for (int q = 0; q < 4; ++q) r.jc[q] -= nj[q];
  } else if (prev != cell) {
    T qc[4], qp[4];
    face_jac4(ro, rd, cur, cr, site(s, prev), radius_of(s, prev), fn, qc, qp);
    for (int q = 0; q < 4; ++q) { r.jc[q] -= qc[q]; r.jp[q] -= qp[q]; }
  }
  if (end_sphere) {
    for (int q = 0; q < 4; ++q) r.jc[q] += fj[q];
  } else if (next != cell) {
    T qc[4], qn[4];
    face_jac4(ro, rd, cur, cr, site(s, next), radius_of(s, next), ff, qc, qn);
    for (int q = 0; q < 4; ++q) { r.jc[q] += qc[q]; r.jn[q] += qn[q]; }
  }
  r.dt = end - start;
  if (r.dt > max_dt) {
    // the clamp is locally constant on its saturated branch
    r.dt = max_dt;
    for (int q = 0; q < 4; ++q) r.jp[q] = r.jc[q] = r.jn[q] = T(0);
  }
  r.tangent = dot4(r.jp, site(s, prev) - ro, radius_of(s, prev)) + dot4(r.jc, cur - ro, cr) +
              dot4(r.jn, site(s, next) - ro, radius_of(s, next));
  *out = r;
  return true;
}

// trace_powerfoam_splats (trace.rs:1456-1546).
template <class T>
TraceOut<T> trace_splats_one(const Scene<T>& s, V3<T> ro, V3<T> dir_in, uint32_t start, const Settings<T>& st,
                             std::vector<SplatInterval<T>>* scratch) {
  TraceOut<T> r{};
  r.last_point = start;
  r.peak_point = 0xFFFFFFFFu;
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return r;
  sorted_splats(s, ro, dir, st.depth, st.max_steps, scratch);
  T trans = 1, ar = 0, ag = 0, ab = 0, depth_mode = 0, peak = 0, t_end = 0;
  uint32_t peak_point = 0xFFFFFFFFu, last = start, steps = 0;
  for (const SplatInterval<T>& iv : *scratch) {
    if (trans <= st.weight_threshold) break;
    steps += 1;
    last = iv.cell;
    t_end = iv.near_t + iv.dt;
    const T sd = s.density[iv.cell];
    if (sd <= T(1e-6f)) continue;
    const T alpha = T(1) - exp_t(-sd * iv.dt);
    const T w = trans * alpha;
    if (st.eval_mode == 1) {
      ar += w * st.const_rgb[0]; ag += w * st.const_rgb[1]; ab += w * st.const_rgb[2];
    } else if (st.eval_mode == 2) {
      V3<T> c = eval_sh_unclamped(s, iv.cell, dir);
      c.x = c.x > T(0) ? c.x : T(0); c.y = c.y > T(0) ? c.y : T(0); c.z = c.z > T(0) ? c.z : T(0);
      ar += w * c.x; ag += w * c.y; ab += w * c.z;
    }
    if (w > peak) {
      peak = w;
      depth_mode = iv.near_t + T(0.5) * iv.dt;
      peak_point = iv.cell;
    }
    trans *= T(1) - alpha;
  }
  r.rgba[0] = ar; r.rgba[1] = ag; r.rgba[2] = ab; r.rgba[3] = T(1) - trans;
  r.steps = steps;
  r.last_point = last;
  r.t_end = t_end;
  r.depth_mode = depth_mode;
  r.peak_weight = peak;
  r.peak_point = peak_point;
  return r;
}

// ---------------------------------------------------------------- point-light shading
// PointLight::sample (relight.rs:203-233), point_light_specular (:249-293) and
// shade_point_light (:295-324).
template <class T> struct Light {
  T pos[3];
  T exponent;
  T dir[3];
  T intensity[3];
  bool visibility = false;  // this light casts a visibility ray even when the lighting's flag is off
};

template <class T> inline bool try_normalize(V3<T>* v) {
  const T len = sqrt_t(dot(*v, *v));
  if (!(len > T(0)) || !std::isfinite(len)) return false;
  const T inv = T(1) / len;  // glam normalize multiplies by the reciprocal length
  *v = inv * (*v);
  return true;
}

template <class T>
void shade_point_light(const Light<T>& L, V3<T> point, V3<T> normal, V3<T> view, const T albedo[3], T roughness,
                       const T f0[3], T out[3]) {
  out[0] = out[1] = out[2] = T(0);
  const V3<T> pos = v3<T>(L.pos[0], L.pos[1], L.pos[2]);
  const V3<T> from = point - pos;
  const T dsq = dot(from, from);
  const T eps = T(1.1920929e-7f);
  if (!std::isfinite(dsq) || dsq <= eps || !(L.exponent >= T(0)) || !std::isfinite(L.exponent)) return;
  for (int c = 0; c < 3; ++c)
    if (!(L.intensity[c] >= T(0)) || !std::isfinite(L.intensity[c])) return;
  const V3<T> away = from / sqrt_t(dsq);
  T angular = T(1);
  if (L.exponent != T(0)) {
    V3<T> d = v3<T>(L.dir[0], L.dir[1], L.dir[2]);
    if (!try_normalize(&d)) return;
    T c = dot(d, away);
    c = c > T(0) ? c : T(0);
    angular = std::pow(c, L.exponent);
  }
  const V3<T> towards = v3<T>(-away.x, -away.y, -away.z);
  T radiance[3];
  for (int c = 0; c < 3; ++c) radiance[c] = L.intensity[c] * angular / dsq;
  V3<T> n = normal;
  T cosine = T(0);
  const bool n_ok = try_normalize(&n);
  if (n_ok) {
    cosine = dot(n, towards);
    cosine = cosine > T(0) ? cosine : T(0);
  }
  for (int c = 0; c < 3; ++c) out[c] = albedo[c] * radiance[c] * cosine;
  if (f0[0] == T(0) && f0[1] == T(0) && f0[2] == T(0)) return;
  V3<T> v = view;
  if (!n_ok || !try_normalize(&v)) return;
  T n_dot_l = dot(n, towards);
  n_dot_l = n_dot_l > T(0) ? n_dot_l : T(0);
  T n_dot_v = dot(n, v);
  n_dot_v = n_dot_v > T(0) ? n_dot_v : T(0);
  V3<T> h = towards + v;
  if (!try_normalize(&h)) return;
  if (n_dot_l == T(0) || n_dot_v == T(0)) return;
  T n_dot_h = dot(n, h);
  n_dot_h = n_dot_h > T(0) ? n_dot_h : T(0);
  T v_dot_h = dot(v, h);
  v_dot_h = v_dot_h < T(0) ? T(0) : (v_dot_h > T(1) ? T(1) : v_dot_h);
  T r = roughness < T(0.045f) ? T(0.045f) : (roughness > T(1) ? T(1) : roughness);
  const T alpha = r * r;
  const T a2 = alpha * alpha;
  const T den = n_dot_h * n_dot_h * (a2 - T(1)) + T(1);
  const T pi = T(3.14159265358979323846);
  T dd = pi * den * den;
  const T tiny = T(1.17549435e-38f);
  dd = dd > tiny ? dd : tiny;
  const T distribution = a2 / dd;
  auto smith = [&](T c) { return T(2) * c / (c + sqrt_t(a2 + (T(1) - a2) * c * c)); };
  const T geometry = smith(n_dot_l) * smith(n_dot_v);
  const T g1 = T(1) - v_dot_h;
  const T grazing = g1 * g1 * g1 * g1 * g1;
  const T scale = pi * distribution * geometry / (T(4) * n_dot_v);
  for (int c = 0; c < 3; ++c) out[c] += radiance[c] * scale * (f0[c] + (T(1) - f0[c]) * grazing);
}

// Relit walk: trace_one_ray's walk and compositing, with each integrated
// segment shaded by every light at the segment's entry point x, normal
// x - site, view -d and the cell's material; light l's sum is scaled by the
// transmittance V_l of an opacity walk from the ray's depth point toward the
// light (start cell = the peak cell) when visibility is on.
template <class T>
void trace_relight(const Scene<T>& s, const T* material, V3<T> ro, V3<T> dir_in, uint32_t start, const Settings<T>& st,
                   const Light<T>* lights, int nl, bool visibility, const T f0[3], T rgba[4]) {
  rgba[0] = rgba[1] = rgba[2] = rgba[3] = T(0);
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return;
  T sums[8][3] = {};
  T t0 = 0, trans = 1, depth_mode = 0, peak = 0;
  uint32_t peak_point = 0xFFFFFFFFu;
  uint32_t cur = start;
  V3<T> cur_pos = site(s, cur);
  T cur_r = radius_of(s, cur);
  const bool bounded = s.radius != nullptr;
  const V3<T> view = v3<T>(-dir.x, -dir.y, -dir.z);
  uint32_t steps = 0;
  while (steps < st.max_steps) {
    steps += 1;
    if (trans <= st.weight_threshold || t0 > st.depth) break;
    T best_t1;
    int64_t face;
    find_exit_face(s, cur, cur_pos, cur_r, ro, dir, t0, st.depth, &best_t1, &face);
    T seg0, seg1;
    if (support_interval(bounded, ro, dir, cur_pos, cur_r, t0, best_t1, &seg0, &seg1)) {
      const T sd = s.density[cur];
      if (sd > T(1e-6f)) {
        const T dt = seg1 - seg0;
        const T alpha = T(1) - exp_t(-sd * dt);
        const T w = trans * alpha;
        const V3<T> x = ro + seg0 * dir;
        const V3<T> nrm = x - cur_pos;
        const T* m = material + 4 * (size_t)cur;
        for (int l = 0; l < nl; ++l) {
          T c[3];
          shade_point_light(lights[l], x, nrm, view, m, m[3], f0, c);
          for (int k = 0; k < 3; ++k) sums[l][k] += w * c[k];
        }
        if (w > peak) {
          peak = w;
          depth_mode = T(0.5) * (seg0 + seg1);
          peak_point = cur;
        }
        trans *= T(1) - alpha;
      }
    }
    if (face < 0) break;
    const uint32_t nb = s.neighbors[s.offsets[cur] + (uint32_t)face];
    t0 = t0 > best_t1 ? t0 : best_t1;
    cur = nb;
    cur_pos = site(s, nb);
    cur_r = radius_of(s, nb);
  }
  T vis[8];
  for (int l = 0; l < nl; ++l) vis[l] = T(1);
  if (peak_point != 0xFFFFFFFFu) {
    const V3<T> p = ro + depth_mode * dir;
    for (int l = 0; l < nl; ++l) {
      if (!visibility && !lights[l].visibility) continue;
      const V3<T> to = v3<T>(lights[l].pos[0], lights[l].pos[1], lights[l].pos[2]) - p;
      const T dist = sqrt_t(dot(to, to));
      if (!(dist > T(0))) continue;
      Settings<T> ss = st;
      ss.depth = dist;
      ss.eval_mode = 0;
      const TraceOut<T> sh = trace_one_ray(s, p, to, peak_point, ss);
      vis[l] = T(1) - sh.rgba[3];
    }
  }
  for (int l = 0; l < nl; ++l)
    for (int k = 0; k < 3; ++k) rgba[k] += vis[l] * sums[l][k];
  rgba[3] = T(1) - trans;
}

// Relit splats: trace_powerfoam_splats' interval list and compositing, each
// composited interval shaded by every light at its entry point x = o + near d,
// normal x - site, view -d and the cell's material; light l's sum is scaled by
// the transmittance of a splat opacity trace from the depth point toward the
// light when visibility is on.
template <class T>
void trace_splat_relight(const Scene<T>& s, const T* material, V3<T> ro, V3<T> dir_in, const Settings<T>& st,
                         const Light<T>* lights, int nl, bool visibility, const T f0[3], T rgba[4],
                         std::vector<SplatInterval<T>>* scratch) {
  rgba[0] = rgba[1] = rgba[2] = rgba[3] = T(0);
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return;
  sorted_splats(s, ro, dir, st.depth, st.max_steps, scratch);
  T sums[8][3] = {};
  T trans = 1, depth_mode = 0, peak = 0;
  uint32_t peak_point = 0xFFFFFFFFu;
  const V3<T> view = v3<T>(-dir.x, -dir.y, -dir.z);
  for (const SplatInterval<T>& iv : *scratch) {
    if (trans <= st.weight_threshold) break;
    const T sd = s.density[iv.cell];
    if (sd <= T(1e-6f)) continue;
    const T alpha = T(1) - exp_t(-sd * iv.dt);
    const T w = trans * alpha;
    const V3<T> x = ro + iv.near_t * dir;
    const V3<T> nrm = x - site(s, iv.cell);
    const T* m = material + 4 * (size_t)iv.cell;
    for (int l = 0; l < nl; ++l) {
      T c[3];
      shade_point_light(lights[l], x, nrm, view, m, m[3], f0, c);
      for (int k = 0; k < 3; ++k) sums[l][k] += w * c[k];
    }
    if (w > peak) {
      peak = w;
      depth_mode = iv.near_t + T(0.5) * iv.dt;
      peak_point = iv.cell;
    }
    trans *= T(1) - alpha;
  }
  T vis[8];
  for (int l = 0; l < nl; ++l) vis[l] = T(1);
  if (peak_point != 0xFFFFFFFFu) {
    const V3<T> p = ro + depth_mode * dir;
    std::vector<SplatInterval<T>> shadow;
    for (int l = 0; l < nl; ++l) {
      if (!visibility && !lights[l].visibility) continue;
      const V3<T> to = v3<T>(lights[l].pos[0], lights[l].pos[1], lights[l].pos[2]) - p;
      const T dist = sqrt_t(dot(to, to));
      if (!(dist > T(0))) continue;
      Settings<T> ss = st;
      ss.depth = dist;
      ss.eval_mode = 0;
      const TraceOut<T> sh = trace_splats_one(s, p, to, 0u, ss, &shadow);
      vis[l] = T(1) - sh.rgba[3];
    }
  }
  for (int l = 0; l < nl; ++l)
    for (int k = 0; k < 3; ++k) rgba[k] += vis[l] * sums[l][k];
  rgba[3] = T(1) - trans;
}

// ---------------------------------------------------------------- Gaussians
// trace_gaussians (trace.rs:695-773): every particle's maximum-response depth
// t* = -(o'.d') / |d'|^2 in its local frame o' = R^-1 (o - mu) / s,
// d' = R^-1 d / s; alpha = opacity exp(-|o' + t* d'|^2 / 2); hits with
// opacity > min_opacity, t* in (t_start, t_end) and alpha >= min_opacity are
// sorted by (t*, index) and composited until T <= min_transmittance, colour =
// max(0, 0.5 + SH(d)).  R^-1 is glam's Quat::inverse (the conjugate) and the
// rotation is glam's Quat * Vec3: v (w^2 - b.b) + b (2 (v.b)) + (b x v) (2 w).
template <class T> struct Gaussians {
  int64_t n;
  const T* mean;      // [n][3]
  const T* opacity;   // [n]
  const T* rot;       // [n][4] x y z w
  const T* scale;     // [n][3]
  const T* sh;        // [n][3K]
  int32_t sh_degree;
};

template <class T> struct GaussSettings {
  T min_opacity, min_transmittance, t_start, t_end;
};

template <class T> inline V3<T> quat_rotate(T qx, T qy, T qz, T qw, V3<T> v) {
  const V3<T> b = v3<T>(qx, qy, qz);
  const T b2 = dot(b, b);
  const T vb = dot(v, b);
  const V3<T> cr = cross(b, v);
  const T s0 = qw * qw - b2, s1 = vb * T(2), s2 = qw * T(2);
  return v3<T>((v.x * s0 + b.x * s1) + cr.x * s2, (v.y * s0 + b.y * s1) + cr.y * s2,
               (v.z * s0 + b.z * s1) + cr.z * s2);
}

template <class T> struct GaussHit {
  T t, alpha;
  uint32_t idx;
};

template <class T>
inline bool gaussian_response(const Gaussians<T>& g, int64_t i, V3<T> ro, V3<T> dir, const GaussSettings<T>& st,
                              GaussHit<T>* out) {
  const T op = g.opacity[i];
  if (op <= st.min_opacity) return false;
  const T* q = g.rot + 4 * i;
  const T* sc = g.scale + 3 * i;
  const V3<T> mu = v3<T>(g.mean[3 * i], g.mean[3 * i + 1], g.mean[3 * i + 2]);
  const V3<T> po = quat_rotate(-q[0], -q[1], -q[2], q[3], ro - mu);
  const V3<T> pd = quat_rotate(-q[0], -q[1], -q[2], q[3], dir);
  const V3<T> lo = v3<T>(po.x / sc[0], po.y / sc[1], po.z / sc[2]);
  const V3<T> ld = v3<T>(pd.x / sc[0], pd.y / sc[1], pd.z / sc[2]);
  const T den = dot(ld, ld);
  const T t = -dot(lo, ld) / den;
  if (!(t > st.t_start && t < st.t_end)) return false;
  const V3<T> c = lo + t * ld;
  const T alpha = op * exp_t(T(-0.5) * dot(c, c));
  if (!(alpha >= st.min_opacity)) return false;
  out->t = t;
  out->alpha = alpha;
  out->idx = (uint32_t)i;
  return true;
}

template <class T>
void sorted_gaussian_hits(const Gaussians<T>& g, V3<T> ro, V3<T> dir, const GaussSettings<T>& st,
                          std::vector<GaussHit<T>>* hits) {
  hits->clear();
  for (int64_t i = 0; i < g.n; ++i) {
    GaussHit<T> h;
    if (gaussian_response(g, i, ro, dir, st, &h)) hits->push_back(h);
  }
  std::sort(hits->begin(), hits->end(), [](const GaussHit<T>& a, const GaussHit<T>& b) {
    return a.t < b.t || (a.t == b.t && a.idx < b.idx);
  });
}

template <class T> inline V3<T> gaussian_colour(const Gaussians<T>& g, uint32_t idx, V3<T> dir) {
  const Scene<T> sc{g.n, nullptr, nullptr, nullptr, g.sh, g.sh_degree, nullptr, nullptr};
  V3<T> c = eval_sh_unclamped(sc, idx, dir);
  c.x = c.x > T(0) ? c.x : T(0);
  c.y = c.y > T(0) ? c.y : T(0);
  c.z = c.z > T(0) ? c.z : T(0);
  return c;
}

// Returns the number of composited hits; rgba = (radiance, 1 - T).
template <class T>
uint32_t trace_gaussians_one(const Gaussians<T>& g, V3<T> ro, V3<T> dir_in, const GaussSettings<T>& st, T rgba[4],
                             std::vector<GaussHit<T>>* hits, uint32_t* order, uint32_t max_order) {
  rgba[0] = rgba[1] = rgba[2] = rgba[3] = T(0);
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return 0;
  sorted_gaussian_hits(g, ro, dir, st, hits);
  T trans = 1, r = 0, gg = 0, b = 0;
  uint32_t count = 0;
  for (const GaussHit<T>& h : *hits) {
    if (trans <= st.min_transmittance) break;
    const V3<T> c = gaussian_colour(g, h.idx, dir);
    const T w = h.alpha * trans;
    r += w * c.x;
    gg += w * c.y;
    b += w * c.z;
    trans *= T(1) - h.alpha;
    if (order && count < max_order) order[count] = h.idx;
    count += 1;
  }
  rgba[0] = r; rgba[1] = gg; rgba[2] = b; rgba[3] = T(1) - trans;
  return count;
}

// Adjoint of trace_gaussians_one given dL/d(rgba): the hit set and order are
// held fixed; gradients are ADDED to d_mean[n][3], d_scale[n][3], d_rot[n][4],
// d_opacity[n], d_sh[n][3K].  With C the final colour and C_<=i the sum over
// the first i hits: dL/dalpha_i = g.(T_i c_i) - g.(C - C_<=i) / (1 - alpha_i)
// + g_a T_final / (1 - alpha_i); alpha = o exp(-q/2) with q = |o' + t* d'|^2
// stationary in t*, so dq/do' = 2 p and dq/dd' = 2 t* p (p the closest point).
template <class T>
void adjoint_gaussians_one(const Gaussians<T>& g, V3<T> ro, V3<T> dir_in, const GaussSettings<T>& st,
                           const T d_rgba[4], T* d_mean, T* d_scale, T* d_rot, T* d_opacity, T* d_sh,
                           std::vector<GaussHit<T>>* hits) {
  V3<T> dir = dir_in;
  if (!normalise(&dir)) return;
  sorted_gaussian_hits(g, ro, dir, st, hits);
  // forward totals
  T trans = 1;
  V3<T> C = v3<T>(0, 0, 0);
  size_t used = 0;
  for (const GaussHit<T>& h : *hits) {
    if (trans <= st.min_transmittance) break;
    const V3<T> c = gaussian_colour(g, h.idx, dir);
    C = C + (h.alpha * trans) * c;
    trans *= T(1) - h.alpha;
    used += 1;
  }
  const T t_final = trans;
  const int K = sh_components(g.sh_degree < 3 ? g.sh_degree : 3);
  T basis[16];
  sh_basis(dir, K, basis);
  const V3<T> gc = v3<T>(d_rgba[0], d_rgba[1], d_rgba[2]);
  const T ga = d_rgba[3];
  T Ti = 1;
  V3<T> acc = v3<T>(0, 0, 0);
  for (size_t k = 0; k < used; ++k) {
    const GaussHit<T>& h = (*hits)[k];
    const int64_t i = h.idx;
    const V3<T> c = gaussian_colour(g, h.idx, dir);
    const T a = h.alpha;
    acc = acc + (a * Ti) * c;
    const V3<T> rest = C - acc;
    const T dA = dot(gc, Ti * c) - dot(gc, rest) / (T(1) - a) + ga * t_final / (T(1) - a);
    // colour: clamp passes where the unclamped value is positive
    const Scene<T> sc{g.n, nullptr, nullptr, nullptr, g.sh, g.sh_degree, nullptr, nullptr};
    const V3<T> cu = eval_sh_unclamped(sc, h.idx, dir);
    const T wgt = a * Ti;
    const T gcol[3] = {cu.x > T(0) ? gc.x * wgt : T(0), cu.y > T(0) ? gc.y * wgt : T(0),
                       cu.z > T(0) ? gc.z * wgt : T(0)};
    for (int kk = 0; kk < K; ++kk)
      for (int ch = 0; ch < 3; ++ch) d_sh[(size_t)i * 3 * K + 3 * kk + ch] += gcol[ch] * basis[kk];
    // alpha = o exp(-q/2)
    const T op = g.opacity[i];
    d_opacity[i] += dA * (a / op);
    const T dq = dA * a * T(-0.5);
    // recompute the local frame
    const T* q = g.rot + 4 * i;
    const T* s = g.scale + 3 * i;
    const V3<T> mu = v3<T>(g.mean[3 * i], g.mean[3 * i + 1], g.mean[3 * i + 2]);
    const V3<T> u = ro - mu;
    const V3<T> bq = v3<T>(-q[0], -q[1], -q[2]);
    const T w = q[3];
    const V3<T> po = quat_rotate(bq.x, bq.y, bq.z, w, u);
    const V3<T> pd = quat_rotate(bq.x, bq.y, bq.z, w, dir);
    const V3<T> lo = v3<T>(po.x / s[0], po.y / s[1], po.z / s[2]);
    const V3<T> ld = v3<T>(pd.x / s[0], pd.y / s[1], pd.z / s[2]);
    const T tt = -dot(lo, ld) / dot(ld, ld);
    const V3<T> p = lo + tt * ld;
    const V3<T> g_lo = (T(2) * dq) * p;
    const V3<T> g_ld = (T(2) * dq * tt) * p;
    const T glo[3] = {g_lo.x, g_lo.y, g_lo.z}, gld[3] = {g_ld.x, g_ld.y, g_ld.z};
    const T lov[3] = {lo.x, lo.y, lo.z}, ldv[3] = {ld.x, ld.y, ld.z};
    T gpo[3], gpd[3];
    for (int a3 = 0; a3 < 3; ++a3) {
      d_scale[3 * i + a3] += -(glo[a3] * lov[a3] + gld[a3] * ldv[a3]) / s[a3];
      gpo[a3] = glo[a3] / s[a3];
      gpd[a3] = gld[a3] / s[a3];
    }
    // f(b, w, v) = v (w^2 - b.b) + b (2 v.b) + (b x v) (2 w); b = -q.xyz
    auto back = [&](V3<T> v, const T gf[3], T gb[3], T* gw, T gv[3]) {
      const T bb[3] = {bq.x, bq.y, bq.z}, vv[3] = {v.x, v.y, v.z};
      const T vb = dot(v, bq);
      const T s0 = w * w - dot(bq, bq);
      // d f_a / d v_c = s0 delta_ac + 2 b_a b_c + 2 w eps(a, ., .): (b x v)_a = eps_abc b_b v_c
      // d f_a / d b_c = -2 b_c v_a + 2 delta_ac vb + 2 b_a v_c + 2 w (e_c x v)_a
      // d f_a / d w   = 2 w v_a + 2 (b x v)_a
      const V3<T> bxv = cross(bq, v);
      const T bxvv[3] = {bxv.x, bxv.y, bxv.z};
      for (int c3 = 0; c3 < 3; ++c3) {
        gv[c3] = 0;
        gb[c3] = 0;
      }
      *gw = 0;
      for (int a3 = 0; a3 < 3; ++a3) {
        *gw += gf[a3] * (T(2) * w * vv[a3] + T(2) * bxvv[a3]);
        for (int c3 = 0; c3 < 3; ++c3) {
          // (b x v)_a derivative wrt v_c: eps(a, b, c) b_b ; wrt b_c: eps(a, c, d) v_d
          T eps_v = 0, eps_b = 0;
          const int a1 = (a3 + 1) % 3, a2 = (a3 + 2) % 3;
          if (c3 == a2) { eps_v = bb[a1]; eps_b = -vv[a1]; }
          if (c3 == a1) { eps_v = -bb[a2]; eps_b = vv[a2]; }
          const T dv = (a3 == c3 ? s0 : T(0)) + T(2) * bb[a3] * bb[c3] + T(2) * w * eps_v;
          const T db = T(-2) * bb[c3] * vv[a3] + (a3 == c3 ? T(2) * vb : T(0)) + T(2) * bb[a3] * vv[c3] +
                       T(2) * w * eps_b;
          gv[c3] += gf[a3] * dv;
          gb[c3] += gf[a3] * db;
        }
      }
    };
    T gb1[3], gw1, gu[3], gb2[3], gw2, gdir[3];
    back(u, gpo, gb1, &gw1, gu);
    back(dir, gpd, gb2, &gw2, gdir);
    for (int a3 = 0; a3 < 3; ++a3) {
      d_mean[3 * i + a3] += -gu[a3];
      d_rot[4 * i + a3] += -(gb1[a3] + gb2[a3]);  // b = -q.xyz
    }
    d_rot[4 * i + 3] += gw1 + gw2;
    Ti *= T(1) - a;
  }
}

// Camera ray of radfoam.wgsl:45-52 with qrot of common.wgsl:19-23.  tan_half is
// computed once on the host (tanf) so device and oracle share it bit for bit.
template <class T>
void camera_ray(const T pos[3], const T quat[4], const T tan_half[2], const T principal[2], uint32_t W, uint32_t H,
                uint32_t x, uint32_t y, T o[3], T d[3]) {
  const T px = (T(x) + T(0.5)) / T(W);
  const T py = (T(y) + T(0.5)) / T(H);
  const T nx = px * T(2) - T(1), ny = py * T(2) - T(1);
  const V3<T> v = v3<T>((nx - principal[0]) * tan_half[0], (ny - principal[1]) * tan_half[1], T(1));
  const V3<T> q = v3<T>(quat[0], quat[1], quat[2]);
  const V3<T> t = T(2) * cross(q, v);
  const V3<T> r = (v + quat[3] * t) + cross(q, t);
  const T len = sqrt_t(dot(r, r));
  o[0] = pos[0]; o[1] = pos[1]; o[2] = pos[2];
  d[0] = r.x / len; d[1] = r.y / len; d[2] = r.z / len;
}

// The batched path recorder's ray of one output row: record_paths reads pixel
// p = pixel_indices[row] + image_pixel_offset (u32 wrap) and walks from the
// camera position along ray_dir_for_pixel(p) (radfoam_record_paths.wgsl:133-153,
// :1227-1230): ix = p % W, iy = p / W, local = ((ndc - principal) * tan_half,
// 1) and the recorder's own rotation qrot(q, v) = 2 (u.v) u + (s^2 - u.u) v +
// 2 s (u x v), evaluated left to right.  The direction is returned before
// normalisation: the walk divides it by its length (the shader's normalize).
inline void recorder_ray(const float pos[3], const float quat[4], const float tan_half[2], const float principal[2],
                         uint32_t W, uint32_t H, uint32_t p, float o[3], float d[3]) {
  const uint32_t ix = p % W, iy = p / W;
  const float px = ((float)ix + 0.5f) / (float)W;
  const float py = ((float)iy + 0.5f) / (float)H;
  const float nx = px * 2.0f - 1.0f, ny = py * 2.0f - 1.0f;
  const V3<float> v = v3<float>((nx - principal[0]) * tan_half[0], (ny - principal[1]) * tan_half[1], 1.0f);
  const V3<float> u = v3<float>(quat[0], quat[1], quat[2]);
  const float s = quat[3];
  const V3<float> r = ((2.0f * dot(u, v)) * u + (s * s - dot(u, u)) * v) + (2.0f * s) * cross(u, v);
  o[0] = pos[0]; o[1] = pos[1]; o[2] = pos[2];
  d[0] = r.x; d[1] = r.y; d[2] = r.z;
}

// containing_cell, lib.rs:526-541: argmin |p - x|^2 - r^2, first index on ties.
template <class T> uint32_t containing_cell(const Scene<T>& s, const T q[3]) {
  uint32_t best = 0;
  T bd = std::numeric_limits<T>::infinity();
  for (int64_t i = 0; i < s.n; ++i) {
    const V3<T> d = v3<T>(q[0], q[1], q[2]) - site(s, (uint32_t)i);
    const T r = radius_of(s, (uint32_t)i);
    const T dist = dot(d, d) - r * r;
    if (dist < bd) { bd = dist; best = (uint32_t)i; }
  }
  return best;
}

template <class T> inline void set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

}  // namespace

extern "C" {

typedef struct {
  int64_t n;
  const float* xyz;
  const float* radius;
  const float* density;
  const float* sh;
  int32_t sh_degree;
  const uint32_t* offsets;
  const uint32_t* neighbors;
} bvo_scene_f32;

typedef struct {
  int64_t n;
  const double* xyz;
  const double* radius;
  const double* density;
  const double* sh;
  int32_t sh_degree;
  const uint32_t* offsets;
  const uint32_t* neighbors;
} bvo_scene_f64;

typedef struct {
  float weight_threshold;
  uint32_t max_steps;
  float depth;
  uint32_t eval_mode;
  float const_rgb[3];
} bvo_settings_f32;

typedef struct {
  double weight_threshold;
  uint32_t max_steps;
  double depth;
  uint32_t eval_mode;
  double const_rgb[3];
} bvo_settings_f64;

}  // extern "C"

namespace {
template <class S, class T> Scene<T> to_scene(const S* d) {
  return Scene<T>{d->n, d->xyz, d->radius, d->density, d->sh, d->sh_degree, d->offsets, d->neighbors};
}
template <class ST, class T> Settings<T> to_settings(const ST* d) {
  return Settings<T>{d->weight_threshold, d->max_steps, d->depth, d->eval_mode,
                     {d->const_rgb[0], d->const_rgb[1], d->const_rgb[2]}};
}

template <class T>
void trace_rays(const Scene<T>& s, int64_t n, const T* o, const T* d, const uint32_t* start, const Settings<T>& st,
                T* rgba, uint32_t* steps, uint32_t* last, T* t_end, T* depth_mode, T* peak_w, uint32_t* peak_pt,
                int threads) {
  set_threads<T>(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    const TraceOut<T> r = trace_one_ray(s, v3<T>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                        v3<T>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), start[i], st);
    for (int c = 0; c < 4; ++c) rgba[4 * i + c] = r.rgba[c];
    if (steps) steps[i] = r.steps;
    if (last) last[i] = r.last_point;
    if (t_end) t_end[i] = r.t_end;
    if (depth_mode) depth_mode[i] = r.depth_mode;
    if (peak_w) peak_w[i] = r.peak_weight;
    if (peak_pt) peak_pt[i] = r.peak_point;
  }
}

// Multi-object scene, trace_scene (scene_trace_core.wgsl:406-465): bounding
// spheres by ray_sphere_intersect (:14-41, t_near clamped to 0), visited in
// (t_near, index) order (object_hit_less, :49-52), each object walked in
// object space from its start cell with integration inside [t_near, t_far]
// (scene_trace_object, :378-402), composited front to back until the
// transmittance drops below weight_threshold.  World directions are
// normalised first, as trace_one_ray does.
template <class T> struct ObjectT {
  Scene<T> s;
  const T* m;       // [12] world -> object, row-major 3x4
  const T* sphere;  // [4] centre xyz, radius
  uint32_t start;
};

template <class T> void trace_objects_one(const std::vector<ObjectT<T>>& objs, V3<T> ro, V3<T> d,
                                          const Settings<T>& st, T out[4]) {
  T cr = 0, cg = 0, cb = 0, tr = 1;
  out[0] = out[1] = out[2] = out[3] = T(0);
  if (!normalise(&d)) return;
  struct Hit { T tn, tf; uint32_t k; };
  std::vector<Hit> hits;
  for (uint32_t k = 0; k < objs.size(); ++k) {
    const T* c = objs[k].sphere;
    const V3<T> oc = ro - v3<T>(c[0], c[1], c[2]);
    const T a = dot(d, d);
    const T b = T(2) * dot(oc, d);
    const T cc = dot(oc, oc) - c[3] * c[3];
    const T disc = b * b - T(4) * a * cc;
    if (disc < T(0)) continue;
    const T sq = std::sqrt(disc);
    const T t1 = (-b - sq) / (T(2) * a);
    const T t2 = (-b + sq) / (T(2) * a);
    if (t2 < T(0)) continue;
    hits.push_back(Hit{t1 > T(0) ? t1 : T(0), t2, k});
  }
  std::sort(hits.begin(), hits.end(),
            [](const Hit& x, const Hit& y) { return x.tn < y.tn || (x.tn == y.tn && x.k < y.k); });
  for (const Hit& h : hits) {
    if (tr < st.weight_threshold) break;
    const ObjectT<T>& ob = objs[h.k];
    const T* m = ob.m;
    const V3<T> oro = v3<T>(m[0] * ro.x + m[1] * ro.y + m[2] * ro.z + m[3], m[4] * ro.x + m[5] * ro.y + m[6] * ro.z + m[7],
                            m[8] * ro.x + m[9] * ro.y + m[10] * ro.z + m[11]);
    const V3<T> od = v3<T>(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
                           m[8] * d.x + m[9] * d.y + m[10] * d.z);
    Settings<T> so = st;
    so.depth = h.tf;
    const TraceOut<T> r = trace_one_ray(ob.s, oro, od, ob.start, so, true, true, h.tn);
    const T alpha = r.rgba[3];
    cr += tr * r.rgba[0];
    cg += tr * r.rgba[1];
    cb += tr * r.rgba[2];
    tr *= T(1) - alpha;
  }
  out[0] = cr; out[1] = cg; out[2] = cb; out[3] = T(1) - tr;
}

template <class S, class ST, class T>
void trace_objects(const S* scenes, const T* xf, const T* spheres, const uint32_t* starts, uint32_t n_obj, int64_t n,
                   const T* o, const T* d, const ST* st_in, T* rgba, int threads) {
  std::vector<ObjectT<T>> objs;
  for (uint32_t k = 0; k < n_obj; ++k)
    objs.push_back(ObjectT<T>{to_scene<S, T>(&scenes[k]), xf + 12 * k, spheres + 4 * k, starts[k]});
  const Settings<T> st = to_settings<ST, T>(st_in);
  set_threads<T>(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i)
    trace_objects_one(objs, v3<T>(o[3 * i], o[3 * i + 1], o[3 * i + 2]), v3<T>(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                      st, rgba + 4 * i);
}
}  // namespace

extern "C" {

void bvo_trace_rays_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, const uint32_t* start,
                        const bvo_settings_f32* st, float* rgba, uint32_t* steps, uint32_t* last, float* t_end,
                        float* depth_mode, float* peak_w, uint32_t* peak_pt, int threads) {
  trace_rays(to_scene<bvo_scene_f32, float>(sc), n, o, d, start, to_settings<bvo_settings_f32, float>(st), rgba, steps,
             last, t_end, depth_mode, peak_w, peak_pt, threads);
}

// trace_one_ray's visited cells: path [n][L] holds the current cell of each of
// the first L loop iterations (the one that stops included); steps [n].
void bvo_trace_paths_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, const uint32_t* start,
                         const bvo_settings_f32* st, uint32_t L, uint32_t* path, uint32_t* steps, int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  const Settings<float> se = to_settings<bvo_settings_f32, float>(st);
  set_threads<float>(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    const TraceOut<float> r =
        trace_one_ray(s, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]), v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                      start[i], se, false, false, 0.0f, path + (size_t)i * L, L);
    steps[i] = r.steps;
  }
}

void bvo_trace_rays_f64(const bvo_scene_f64* sc, int64_t n, const double* o, const double* d, const uint32_t* start,
                        const bvo_settings_f64* st, double* rgba, uint32_t* steps, uint32_t* last, double* t_end,
                        double* depth_mode, double* peak_w, uint32_t* peak_pt, int threads) {
  trace_rays(to_scene<bvo_scene_f64, double>(sc), n, o, d, start, to_settings<bvo_settings_f64, double>(st), rgba,
             steps, last, t_end, depth_mode, peak_w, peak_pt, threads);
}

void bvo_trace_objects_f32(const bvo_scene_f32* scenes, const float* xf, const float* spheres, const uint32_t* starts,
                           uint32_t n_obj, int64_t n, const float* o, const float* d, const bvo_settings_f32* st,
                           float* rgba, int threads) {
  trace_objects(scenes, xf, spheres, starts, n_obj, n, o, d, st, rgba, threads);
}

void bvo_trace_objects_f64(const bvo_scene_f64* scenes, const double* xf, const double* spheres,
                           const uint32_t* starts, uint32_t n_obj, int64_t n, const double* o, const double* d,
                           const bvo_settings_f64* st, double* rgba, int threads) {
  trace_objects(scenes, xf, spheres, starts, n_obj, n, o, d, st, rgba, threads);
}

void bvo_record_paths_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, const uint32_t* start,
                          uint32_t L, float depth, float max_dt, uint32_t* cells, uint32_t* next_cells, float* dts,
                          uint32_t* status, int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  set_threads<float>(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    status[i] = record_path(s, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                            v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), start[i], L, depth, max_dt,
                            cells + (size_t)i * L, next_cells + (size_t)i * L, dts + (size_t)i * L);
  }
}

// Gradient accumulation is serial over rays so the result is deterministic.
void bvo_adjoint_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, const uint32_t* start,
                     const bvo_settings_f32* st, const float* d_rgba, float* d_density, float* d_sh, float* d_xyz) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  const Settings<float> se = to_settings<bvo_settings_f32, float>(st);
  for (int64_t i = 0; i < n; ++i)
    adjoint_one_ray(s, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]), v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                    start[i], se, d_rgba + 4 * i, d_density, d_sh, d_xyz);
}

// Threads an OpenMP loop with threads = 0 uses (OMP_NUM_THREADS or the cores).
int bvo_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Timing baseline only: the FP32 adjoint over `threads` OpenMP threads (0 =
// all), gradients added atomically, so their summation order varies run to run.
void bvo_adjoint_mt_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, const uint32_t* start,
                        const bvo_settings_f32* st, const float* d_rgba, float* d_density, float* d_sh, float* d_xyz,
                        int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  const Settings<float> se = to_settings<bvo_settings_f32, float>(st);
  set_threads<float>(threads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i)
    adjoint_one_ray<float, true>(s, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                 v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), start[i], se, d_rgba + 4 * i,
                                 d_density, d_sh, d_xyz);
}

void bvo_adjoint_f64(const bvo_scene_f64* sc, int64_t n, const double* o, const double* d, const uint32_t* start,
                     const bvo_settings_f64* st, const double* d_rgba, double* d_density, double* d_sh,
                     double* d_xyz) {
  const Scene<double> s = to_scene<bvo_scene_f64, double>(sc);
  const Settings<double> se = to_settings<bvo_settings_f64, double>(st);
  for (int64_t i = 0; i < n; ++i)
    adjoint_one_ray(s, v3<double>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                    v3<double>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), start[i], se, d_rgba + 4 * i, d_density, d_sh,
                    d_xyz);
}

typedef struct {
  float position[3];
  float exponent;
  float direction[3];
  uint32_t visibility;
  float intensity[3];
  float pad1;
} bvo_point_light;

void bvo_trace_relight_f32(const bvo_scene_f32* sc, const float* material, int64_t n, const float* o, const float* d,
                           const uint32_t* start, const bvo_settings_f32* st, const bvo_point_light* lights, int nl,
                           int visibility, const float f0[3], float* rgba, int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  const Settings<float> se = to_settings<bvo_settings_f32, float>(st);
  Light<float> L[8];
  nl = nl < 8 ? nl : 8;
  for (int l = 0; l < nl; ++l) {
    for (int k = 0; k < 3; ++k) {
      L[l].pos[k] = lights[l].position[k];
      L[l].dir[k] = lights[l].direction[k];
      L[l].intensity[k] = lights[l].intensity[k];
    }
    L[l].exponent = lights[l].exponent;
    L[l].visibility = lights[l].visibility != 0;
  }
  set_threads<float>(threads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i)
    trace_relight(s, material, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                  v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), start[i], se, L, nl, visibility != 0, f0,
                  rgba + 4 * i);
}

// shade_point_light (relight.rs:295) at one point, for direct tests.
void bvo_shade_point_light_f32(const bvo_point_light* light, const float point[3], const float normal[3],
                               const float view[3], const float albedo[3], float roughness, const float f0[3],
                               float out[3]) {
  Light<float> L;
  for (int k = 0; k < 3; ++k) {
    L.pos[k] = light->position[k];
    L.dir[k] = light->direction[k];
    L.intensity[k] = light->intensity[k];
  }
  L.exponent = light->exponent;
  shade_point_light(L, v3<float>(point[0], point[1], point[2]), v3<float>(normal[0], normal[1], normal[2]),
                    v3<float>(view[0], view[1], view[2]), albedo, roughness, f0, out);
}

void bvo_camera_rays_f32(const float pos[3], const float quat[4], const float tan_half[2], const float principal[2],
                         uint32_t W, uint32_t H, float* o, float* d) {
  for (uint32_t y = 0; y < H; ++y)
    for (uint32_t x = 0; x < W; ++x) {
      const size_t i = (size_t)y * W + x;
      camera_ray(pos, quat, tan_half, principal, W, H, x, y, o + 3 * i, d + 3 * i);
    }
}

void bvo_recorder_rays_f32(const float pos[3], const float quat[4], const float tan_half[2], const float principal[2],
                           uint32_t W, uint32_t H, uint32_t image_pixel_offset, const uint32_t* pixel_indices,
                           int64_t n, float* o, float* d) {
  for (int64_t i = 0; i < n; ++i)
    recorder_ray(pos, quat, tan_half, principal, W, H, pixel_indices[i] + image_pixel_offset, o + 3 * i, d + 3 * i);
}

uint32_t bvo_containing_cell_f32(const bvo_scene_f32* sc, const float q[3]) {
  return containing_cell(to_scene<bvo_scene_f32, float>(sc), q);
}

// SH colour (clamped, trace.rs:397-402) of one cell, for direct SH tests.
void bvo_eval_sh_f32(const bvo_scene_f32* sc, uint32_t cell, const float dir[3], float out[3]) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  V3<float> c = eval_sh_unclamped(s, cell, v3<float>(dir[0], dir[1], dir[2]));
  out[0] = c.x > 0.f ? c.x : 0.f;
  out[1] = c.y > 0.f ? c.y : 0.f;
  out[2] = c.z > 0.f ? c.z : 0.f;
}

// support_interval exposed for the closed-form tests of trace.rs:2444-2461.
int bvo_support_interval_f32(int bounded, const float ro[3], const float rd[3], const float c[3], float radius,
                             float t0, float t1, float out[2]) {
  return support_interval<float>(bounded != 0, v3<float>(ro[0], ro[1], ro[2]), v3<float>(rd[0], rd[1], rd[2]),
                                 v3<float>(c[0], c[1], c[2]), radius, t0, t1, &out[0], &out[1])
             ? 1
             : 0;
}

// trace_powerfoam_splats over a batch of rays (outputs as bvo_trace_rays_f32).
void bvo_trace_splats_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, const uint32_t* start,
                          const bvo_settings_f32* st, float* rgba, uint32_t* steps, uint32_t* last, float* t_end,
                          float* depth_mode, float* peak_w, uint32_t* peak_pt, int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  const Settings<float> se = to_settings<bvo_settings_f32, float>(st);
  set_threads<float>(threads);
#pragma omp parallel
  {
    std::vector<SplatInterval<float>> scratch;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
      const TraceOut<float> r = trace_splats_one(s, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                                 v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]),
                                                 start ? start[i] : 0u, se, &scratch);
      for (int c = 0; c < 4; ++c) rgba[4 * i + c] = r.rgba[c];
      if (steps) steps[i] = r.steps;
      if (last) last[i] = r.last_point;
      if (t_end) t_end[i] = r.t_end;
      if (depth_mode) depth_mode[i] = r.depth_mode;
      if (peak_w) peak_w[i] = r.peak_weight;
      if (peak_pt) peak_pt[i] = r.peak_point;
    }
  }
}

// record_powerfoam_splats_jacobians (trace.rs:1414-1443) without the Jacobians:
// row i holds the first L sorted intervals (cell, near, dt); count[i] = how many.
void bvo_record_splats_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, float depth,
                           uint32_t L, uint32_t* cells, float* nears, float* dts, uint32_t* count, int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  set_threads<float>(threads);
#pragma omp parallel
  {
    std::vector<SplatInterval<float>> iv;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
      V3<float> dir = v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
      count[i] = 0;
      if (!normalise(&dir)) continue;
      sorted_splats(s, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]), dir, depth, L, &iv);
      for (size_t k = 0; k < iv.size(); ++k) {
        cells[(size_t)i * L + k] = iv[k].cell;
        nears[(size_t)i * L + k] = iv[k].near_t;
        dts[(size_t)i * L + k] = iv[k].dt;
      }
      count[i] = (uint32_t)iv.size();
    }
  }
}

// The GPU recorder's PowerFoam path with Jacobian streams (record_powerfoam_splats,
// radfoam_record_paths.wgsl, jacobian_mode 1; the CPU counterpart is
// record_powerfoam_splats_jacobians, trace.rs:1414-1443): row i holds the first
// L sorted intervals as cell, next, previous, dt (clamped to max_dt), the three
// [4] Jacobian rows and the reference tangent; status[i] = count | bit 31 when
// more intervals follow.
}  // extern "C"

namespace {
template <class T, class S>
void record_splat_jac(const S* sc, int64_t n, const T* o, const T* d, T depth, uint32_t L, T max_dt, uint32_t* cells,
                      uint32_t* next, uint32_t* prev, T* dts, T* jp, T* jc, T* jn, T* tangent, uint32_t* status,
                      int threads) {
  const Scene<T> s = to_scene<S, T>(sc);
  set_threads<T>(threads);
#pragma omp parallel
  {
    std::vector<std::pair<T, uint32_t>> order;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
      V3<T> dir = v3<T>(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
      status[i] = 0;
      if (!normalise(&dir)) continue;
      const V3<T> ro = v3<T>(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
      order.clear();
      for (int64_t c = 0; c < s.n; ++c) {
        SplatJac<T> iv;
        if (splat_interval_jac(s, ro, dir, depth, (uint32_t)c, max_dt, &iv)) order.push_back({iv.near_t, (uint32_t)c});
      }
      std::sort(order.begin(), order.end(), [](const std::pair<T, uint32_t>& a, const std::pair<T, uint32_t>& b) {
        return a.first < b.first || (a.first == b.first && a.second < b.second);
      });
      const size_t cnt = order.size() < L ? order.size() : L;
      for (size_t k = 0; k < cnt; ++k) {
        SplatJac<T> iv;
        splat_interval_jac(s, ro, dir, depth, order[k].second, max_dt, &iv);
        const size_t r = (size_t)i * L + k;
        cells[r] = iv.cell;
        next[r] = iv.next;
        prev[r] = iv.prev;
        dts[r] = iv.dt;
        tangent[r] = iv.tangent;
        for (int q = 0; q < 4; ++q) {
          jp[4 * r + q] = iv.jp[q];
          jc[4 * r + q] = iv.jc[q];
          jn[4 * r + q] = iv.jn[q];
        }
      }
      status[i] = (uint32_t)cnt | (order.size() > L ? 0x80000000u : 0u);
    }
  }
}
}  // namespace

extern "C" {

void bvo_record_splat_jac_f32(const bvo_scene_f32* sc, int64_t n, const float* o, const float* d, float depth,
                              uint32_t L, float max_dt, uint32_t* cells, uint32_t* next, uint32_t* prev, float* dts,
                              float* jp, float* jc, float* jn, float* tangent, uint32_t* status, int threads) {
  record_splat_jac<float>(sc, n, o, d, depth, L, max_dt, cells, next, prev, dts, jp, jc, jn, tangent, status, threads);
}

// One cell's interval in FP64 (the finite-difference twin): returns 1 and
// fills (near, dt, prev, next, jp, jc, jn) when the ray's interval is non-empty.
int bvo_splat_interval_jac_f64(const bvo_scene_f64* sc, const double o[3], const double d[3], double depth,
                               uint32_t cell, double max_dt, double out[14], uint32_t ids[2]) {
  const Scene<double> s = to_scene<bvo_scene_f64, double>(sc);
  V3<double> dir = v3<double>(d[0], d[1], d[2]);
  if (!normalise(&dir)) return 0;
  SplatJac<double> iv;
  if (!splat_interval_jac(s, v3<double>(o[0], o[1], o[2]), dir, depth, cell, max_dt, &iv)) return 0;
  out[0] = iv.near_t;
  out[1] = iv.dt;
  for (int q = 0; q < 4; ++q) {
    out[2 + q] = iv.jp[q];
    out[6 + q] = iv.jc[q];
    out[10 + q] = iv.jn[q];
  }
  ids[0] = iv.prev;
  ids[1] = iv.next;
  return 1;
}

void bvo_trace_splat_relight_f32(const bvo_scene_f32* sc, const float* material, int64_t n, const float* o,
                                 const float* d, const bvo_settings_f32* st, const bvo_point_light* lights, int nl,
                                 int visibility, const float f0[3], float* rgba, int threads) {
  const Scene<float> s = to_scene<bvo_scene_f32, float>(sc);
  const Settings<float> se = to_settings<bvo_settings_f32, float>(st);
  Light<float> L[8];
  nl = nl < 8 ? nl : 8;
  for (int l = 0; l < nl; ++l) {
    for (int k = 0; k < 3; ++k) {
      L[l].pos[k] = lights[l].position[k];
      L[l].dir[k] = lights[l].direction[k];
      L[l].intensity[k] = lights[l].intensity[k];
    }
    L[l].exponent = lights[l].exponent;
    L[l].visibility = lights[l].visibility != 0;
  }
  set_threads<float>(threads);
#pragma omp parallel
  {
    std::vector<SplatInterval<float>> scratch;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i)
      trace_splat_relight(s, material, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                          v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), se, L, nl, visibility != 0, f0,
                          rgba + 4 * i, &scratch);
  }
}

typedef struct {
  int64_t n;
  const void* mean;
  const void* opacity;
  const void* rot;
  const void* scale;
  const void* sh;
  int32_t sh_degree;
} bvo_gaussians;

}  // extern "C"

namespace {
template <class T> Gaussians<T> to_gauss(const bvo_gaussians* d) {
  return Gaussians<T>{d->n, (const T*)d->mean, (const T*)d->opacity, (const T*)d->rot, (const T*)d->scale,
                      (const T*)d->sh, d->sh_degree};
}
}  // namespace

extern "C" {

// trace_gaussians over a batch; hits[i] = composited count; order rows [n][L]
// (may be null) receive the composited particle indices in order.
void bvo_trace_gaussians_f32(const bvo_gaussians* gd, int64_t n, const float* o, const float* d, const float st4[4],
                             float* rgba, uint32_t* hits, uint32_t* order, uint32_t L, int threads) {
  const Gaussians<float> g = to_gauss<float>(gd);
  const GaussSettings<float> st{st4[0], st4[1], st4[2], st4[3]};
  set_threads<float>(threads);
#pragma omp parallel
  {
    std::vector<GaussHit<float>> scratch;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
      const uint32_t c = trace_gaussians_one(g, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                                             v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), st, rgba + 4 * i,
                                             &scratch, order ? order + (size_t)i * L : nullptr, L);
      if (hits) hits[i] = c;
    }
  }
}

void bvo_trace_gaussians_f64(const bvo_gaussians* gd, int64_t n, const double* o, const double* d,
                             const double st4[4], double* rgba, int threads) {
  const Gaussians<double> g = to_gauss<double>(gd);
  const GaussSettings<double> st{st4[0], st4[1], st4[2], st4[3]};
  set_threads<double>(threads);
#pragma omp parallel
  {
    std::vector<GaussHit<double>> scratch;
#pragma omp for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i)
      trace_gaussians_one(g, v3<double>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                          v3<double>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), st, rgba + 4 * i, &scratch, nullptr, 0);
  }
}

// Serial over rays so the accumulation order is fixed.
void bvo_adjoint_gaussians_f32(const bvo_gaussians* gd, int64_t n, const float* o, const float* d,
                               const float st4[4], const float* d_rgba, float* d_mean, float* d_scale, float* d_rot,
                               float* d_opacity, float* d_sh) {
  const Gaussians<float> g = to_gauss<float>(gd);
  const GaussSettings<float> st{st4[0], st4[1], st4[2], st4[3]};
  std::vector<GaussHit<float>> scratch;
  for (int64_t i = 0; i < n; ++i)
    adjoint_gaussians_one(g, v3<float>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                          v3<float>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), st, d_rgba + 4 * i, d_mean, d_scale, d_rot,
                          d_opacity, d_sh, &scratch);
}

void bvo_adjoint_gaussians_f64(const bvo_gaussians* gd, int64_t n, const double* o, const double* d,
                               const double st4[4], const double* d_rgba, double* d_mean, double* d_scale,
                               double* d_rot, double* d_opacity, double* d_sh) {
  const Gaussians<double> g = to_gauss<double>(gd);
  const GaussSettings<double> st{st4[0], st4[1], st4[2], st4[3]};
  std::vector<GaussHit<double>> scratch;
  for (int64_t i = 0; i < n; ++i)
    adjoint_gaussians_one(g, v3<double>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                          v3<double>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), st, d_rgba + 4 * i, d_mean, d_scale,
                          d_rot, d_opacity, d_sh, &scratch);
}

}  // extern "C"
