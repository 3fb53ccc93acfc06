// Delaunay adjacency of the sites: the edge graph of their 3-D Delaunay
// tetrahedralization as a CSR with sorted, self-free rows — what
// adjacency::try_compute_adjacency (blade-volume/src/adjacency.rs:142-197)
// builds with simple_delaunay_lib, or with Qhull (:564-634).
//
// Incremental Bowyer-Watson insertion over an infinite vertex, host C++:
//   * sites are inserted in a biased randomized order (rounds of doubling size,
//     each round in Morton order), so point location walks are short;
//   * location is a visibility walk from the last created tetrahedron, starting
//     each step at a rotating face;
//   * the cavity is the connected set of tetrahedra whose circumsphere holds
//     the new site; an infinite tetrahedron (a hull face and the vertex at
//     infinity) is in conflict when the site lies strictly beyond its face,
//     and, when the site is coplanar with the face, when it lies inside the
//     circumsphere of the finite tetrahedron behind that face (whose trace on
//     the plane is the face's circumcircle);
//   * orient3d and insphere are exact: a double-precision evaluation with a
//     forward error bound, and expansion arithmetic (Shewchuk, "Adaptive
//     Precision Floating-Point Arithmetic and Fast Robust Geometric
//     Predicates", 1997) when the bound does not decide the sign;
//   * an insphere test that is exactly zero is decided by symbolic
//     perturbation of the lifted coordinate, the site of larger index carrying
//     the larger perturbation, so cospherical sites get one consistent
//     triangulation;
//   * a site equal to an inserted one is skipped and gets an empty row.
// In general position the Delaunay tetrahedralization is unique, so the edge
// set equals any exact builder's (Qhull's included).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "bv_host.h"

namespace {

// ---------------------------------------------------------------- exact predicates
constexpr double kEps = 1.1102230246251565e-16;  // 2^-53
constexpr double kO3dBound = 16.0 * kEps;         // above (7 + 56 eps) eps for this evaluation order
constexpr double kIspBound = 64.0 * kEps;         // above (16 + 224 eps) eps for this evaluation order

inline void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  const double bv = x - a, av = x - bv, br = b - bv, ar = a - av;
  y = ar + br;
}
inline void two_diff(double a, double b, double& x, double& y) {
  x = a - b;
  const double bv = a - x, av = x + bv, br = bv - b, ar = a - av;
  y = ar + br;
}
inline void two_prod(double a, double b, double& x, double& y) {
  x = a * b;
  y = std::fma(a, b, -x);
}

// Nonoverlapping expansions, components in increasing magnitude, zeros elided.
using Exp = std::vector<double>;

Exp grow(const Exp& e, double b) {
  Exp h;
  h.reserve(e.size() + 1);
  double q = b;
  for (double ei : e) {
    double x, y;
    two_sum(q, ei, x, y);
    if (y != 0.0) h.push_back(y);
    q = x;
  }
  if (q != 0.0 || h.empty()) h.push_back(q);
  return h;
}
Exp esum(Exp e, const Exp& f) {
  for (double fi : f) e = grow(e, fi);
  return e;
}
Exp eneg(Exp e) {
  for (double& v : e) v = -v;
  return e;
}
Exp escale(const Exp& e, double b) {
  Exp r;
  for (double ei : e) {
    double p, q;
    two_prod(ei, b, p, q);
    r = grow(grow(r, q), p);
  }
  return r;
}
Exp emul(const Exp& e, const Exp& f) {
  Exp r;
  for (double fj : f) r = esum(r, escale(e, fj));
  return r;
}
int esign(const Exp& e) {
  for (size_t i = e.size(); i-- > 0;)
    if (e[i] != 0.0) return e[i] > 0.0 ? 1 : -1;
  return 0;
}
Exp ediff(double a, double b) {
  double x, y;
  two_diff(a, b, x, y);
  Exp r;
  if (y != 0.0) r.push_back(y);
  r.push_back(x);
  return r;
}

struct E3 {
  Exp x, y, z;
};
E3 diff3(const double* p, const double* q) { return E3{ediff(p[0], q[0]), ediff(p[1], q[1]), ediff(p[2], q[2])}; }
// p . (q x r)
Exp det3(const E3& p, const E3& q, const E3& r) {
  const Exp cx = esum(emul(q.y, r.z), eneg(emul(q.z, r.y)));
  const Exp cy = esum(emul(q.z, r.x), eneg(emul(q.x, r.z)));
  const Exp cz = esum(emul(q.x, r.y), eneg(emul(q.y, r.x)));
  return esum(esum(emul(p.x, cx), emul(p.y, cy)), emul(p.z, cz));
}

int orient_exact(const double* a, const double* b, const double* c, const double* d) {
  return esign(det3(diff3(a, d), diff3(b, d), diff3(c, d)));
}

int insphere_exact(const double* a, const double* b, const double* c, const double* d, const double* e) {
  const E3 r[4] = {diff3(a, e), diff3(b, e), diff3(c, e), diff3(d, e)};
  Exp lift[4];
  for (int i = 0; i < 4; ++i) lift[i] = esum(esum(emul(r[i].x, r[i].x), emul(r[i].y, r[i].y)), emul(r[i].z, r[i].z));
  // expansion along the lift column: sum_i (-1)^(i+3) lift_i det3(other rows)
  const Exp ma = det3(r[1], r[2], r[3]), mb = det3(r[0], r[2], r[3]);
  const Exp mc = det3(r[0], r[1], r[3]), md = det3(r[0], r[1], r[2]);
  Exp s = esum(eneg(emul(lift[0], ma)), emul(lift[1], mb));
  s = esum(s, eneg(emul(lift[2], mc)));
  s = esum(s, emul(lift[3], md));
  return esign(s);
}

// orient(a, b, c, d) = (a - d) . ((b - d) x (c - d)): the sign of the
// tetrahedron's signed volume.
inline int orient(const double* a, const double* b, const double* c, const double* d) {
  const double adx = a[0] - d[0], ady = a[1] - d[1], adz = a[2] - d[2];
  const double bdx = b[0] - d[0], bdy = b[1] - d[1], bdz = b[2] - d[2];
  const double cdx = c[0] - d[0], cdy = c[1] - d[1], cdz = c[2] - d[2];
  const double m1 = bdy * cdz, m2 = bdz * cdy, m3 = bdz * cdx, m4 = bdx * cdz, m5 = bdx * cdy, m6 = bdy * cdx;
  const double det = adx * (m1 - m2) + ady * (m3 - m4) + adz * (m5 - m6);
  const double perm = std::fabs(adx) * (std::fabs(m1) + std::fabs(m2)) +
                      std::fabs(ady) * (std::fabs(m3) + std::fabs(m4)) +
                      std::fabs(adz) * (std::fabs(m5) + std::fabs(m6));
  const double bound = kO3dBound * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return orient_exact(a, b, c, d);
}

inline double det3d(const double* p, const double* q, const double* r, double* perm) {
  const double m1 = q[1] * r[2], m2 = q[2] * r[1], m3 = q[2] * r[0], m4 = q[0] * r[2], m5 = q[0] * r[1],
               m6 = q[1] * r[0];
  *perm = std::fabs(p[0]) * (std::fabs(m1) + std::fabs(m2)) + std::fabs(p[1]) * (std::fabs(m3) + std::fabs(m4)) +
          std::fabs(p[2]) * (std::fabs(m5) + std::fabs(m6));
  return p[0] * (m1 - m2) + p[1] * (m3 - m4) + p[2] * (m5 - m6);
}

// > 0 when e lies inside the circumsphere of the positively oriented (a, b, c, d).
inline int insphere(const double* a, const double* b, const double* c, const double* d, const double* e) {
  double r[4][3], lift[4];
  const double* p[4] = {a, b, c, d};
  for (int i = 0; i < 4; ++i) {
    for (int k = 0; k < 3; ++k) r[i][k] = p[i][k] - e[k];
    lift[i] = r[i][0] * r[i][0] + r[i][1] * r[i][1] + r[i][2] * r[i][2];
  }
  double pa, pb, pc, pd;
  const double ma = det3d(r[1], r[2], r[3], &pa), mb = det3d(r[0], r[2], r[3], &pb);
  const double mc = det3d(r[0], r[1], r[3], &pc), md = det3d(r[0], r[1], r[2], &pd);
  const double det = -lift[0] * ma + lift[1] * mb - lift[2] * mc + lift[3] * md;
  const double perm = lift[0] * pa + lift[1] * pb + lift[2] * pc + lift[3] * pd;
  const double bound = kIspBound * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return insphere_exact(a, b, c, d, e);
}

// ---------------------------------------------------------------- triangulation
constexpr int32_t kInf = -1;

struct Tet {
  int32_t v[4];  // vertices; kInf for the vertex at infinity (at most one)
  int32_t n[4];  // n[i]: the tetrahedron across the face opposite v[i]
};

struct Triangulation {
  const double* P = nullptr;  // [n][3]
  std::vector<Tet> tets;
  std::vector<uint8_t> dead;
  std::vector<int32_t> free_list;
  std::vector<uint32_t> stamp;
  uint32_t round = 0;
  int32_t hint = 0;
  uint32_t rot = 0;
  // scratch of one insertion
  std::vector<int32_t> cavity, stack, created;
  // a cavity face: cavity tetrahedron t, its face i, and the index of the
  // face of the outer tetrahedron that points back at t
  struct Face {
    int32_t t;
    int i, back;
  };
  std::vector<Face> boundary;
  std::vector<Tet> olds;
  struct Slot {
    uint64_t key;
    int32_t tet;
    int face;
  };
  std::vector<Slot> edge_table;
  std::vector<uint32_t> edge_stamp;

  const double* pt(int32_t i) const { return P + 3 * (size_t)i; }

  int32_t alloc() {
    if (!free_list.empty()) {
      const int32_t t = free_list.back();
      free_list.pop_back();
      dead[t] = 0;
      return t;
    }
    tets.push_back(Tet{});
    dead.push_back(0);
    stamp.push_back(0);
    return (int32_t)tets.size() - 1;
  }

  // orient with vertex slot i of t replaced by q (t finite).
  int orient_with(const Tet& t, int i, const double* q) const {
    const double* p[4];
    for (int k = 0; k < 4; ++k) p[k] = k == i ? q : pt(t.v[k]);
    return orient(p[0], p[1], p[2], p[3]);
  }

  // insphere of a finite t with the query site s, ties broken by perturbing
  // the lifted coordinate of each of the five sites by eps^rank, the larger
  // index carrying the larger perturbation.  With the lifted determinant D
  // (rows a, b, c, d, s; columns x, y, z, |p|^2, 1) the perturbed value is
  // D + sum_i delta_i C_i, C_i = (-1)^(i+3) orient(the other four rows).
  bool conflict_finite(const Tet& t, int32_t s) const {
    const double* q = pt(s);
    const int v = insphere(pt(t.v[0]), pt(t.v[1]), pt(t.v[2]), pt(t.v[3]), q);
    if (v != 0) return v > 0;
    int32_t ids[5] = {t.v[0], t.v[1], t.v[2], t.v[3], s};
    int order[5] = {0, 1, 2, 3, 4};
    std::sort(order, order + 5, [&](int x, int y) { return ids[x] > ids[y]; });
    for (int r = 0; r < 5; ++r) {
      const int i = order[r];
      const double* o[4];
      int m = 0;
      for (int k = 0; k < 5; ++k)
        if (k != i) o[m++] = pt(ids[k]);
      const int c = orient(o[0], o[1], o[2], o[3]);
      if (c != 0) return ((i + 3) % 2 == 0 ? c : -c) > 0;
    }
    return false;  // unreachable: t's own four vertices are not coplanar
  }

  bool conflict(int32_t ti, int32_t s) const {
    const Tet& t = tets[ti];
    int k = -1;
    for (int i = 0; i < 4; ++i)
      if (t.v[i] == kInf) k = i;
    if (k < 0) return conflict_finite(t, s);
    const int o = orient_with(t, k, pt(s));
    if (o != 0) return o > 0;
    // coplanar with the hull face: inside its circumcircle, i.e. inside the
    // circumsphere of the finite tetrahedron behind it
    return conflict_finite(tets[t.n[k]], s);
  }

  bool has_inf(const Tet& t) const { return t.v[0] == kInf || t.v[1] == kInf || t.v[2] == kInf || t.v[3] == kInf; }

  int32_t finite_behind(int32_t t) const {
    const Tet& T = tets[t];
    for (int i = 0; i < 4; ++i)
      if (T.v[i] == kInf) return T.n[i];
    return t;
  }

  // A tetrahedron in conflict with s, or -1 when s equals an inserted site
  // (*dup = that site).  The walk runs through finite tetrahedra; stepping
  // through a hull face means s lies strictly beyond it, so the infinite
  // tetrahedron there is in conflict.
  int32_t locate(int32_t s, int32_t* dup) {
    const double* q = pt(s);
    int32_t t = hint;
    if (t < 0 || t >= (int32_t)tets.size() || dead[t]) {
      t = 0;
      while (dead[t]) ++t;
    }
    t = finite_behind(t);
    size_t steps = 0;
    const size_t limit = 64 + 8 * tets.size();
    for (;;) {
      const Tet& T = tets[t];
      if (has_inf(T)) return t;
      int next = -1;
      const uint32_t r0 = rot++;
      for (int j = 0; j < 4; ++j) {
        const int i = (int)((r0 + j) & 3u);
        if (orient_with(T, i, q) < 0) {
          next = i;
          break;
        }
      }
      if (next < 0) {
        for (int i = 0; i < 4; ++i) {
          const double* p = pt(T.v[i]);
          if (p[0] == q[0] && p[1] == q[1] && p[2] == q[2]) {
            *dup = T.v[i];
            return -1;
          }
        }
        return t;
      }
      t = T.n[next];
      if (++steps > limit) break;
    }
    // a walk that does not end (it cannot in a Delaunay triangulation): scan
    for (size_t u = 0; u < tets.size(); ++u)
      if (!dead[u] && conflict((int32_t)u, s)) return (int32_t)u;
    return -1;
  }

  static uint64_t edge_key(int32_t a, int32_t b) {
    const uint64_t x = (uint32_t)a, y = (uint32_t)b;  // kInf maps to 0xFFFFFFFF
    return x < y ? (x << 32 | y) : (y << 32 | x);
  }

  void insert(int32_t s, int32_t* dup) {
    *dup = -1;
    const int32_t t0 = locate(s, dup);
    if (t0 < 0) return;
    ++round;
    // mark[t] = 2 round + 1: in the cavity; 2 round: tested, outside
    const uint32_t in_mark = 2 * round + 1, out_mark = 2 * round;
    cavity.clear();
    boundary.clear();
    stack.clear();
    stack.push_back(t0);
    stamp[t0] = in_mark;
    while (!stack.empty()) {
      const int32_t t = stack.back();
      stack.pop_back();
      cavity.push_back(t);
      for (int i = 0; i < 4; ++i) {
        const int32_t u = tets[t].n[i];
        if (stamp[u] == in_mark) continue;
        if (stamp[u] != out_mark && conflict(u, s)) {
          stamp[u] = in_mark;
          stack.push_back(u);
        } else {
          stamp[u] = out_mark;
          int back = 0;
          while (tets[u].n[back] != t) ++back;
          boundary.push_back(Face{t, i, back});
        }
      }
    }
    // each boundary face joined to s; the cavity's slots are reused, so the
    // old tetrahedra are copied first
    const size_t nb = boundary.size();
    olds.resize(nb);
    for (size_t b = 0; b < nb; ++b) olds[b] = tets[boundary[b].t];
    for (int32_t t : cavity) {
      dead[t] = 1;
      free_list.push_back(t);
    }
    created.resize(nb);
    for (size_t b = 0; b < nb; ++b) created[b] = alloc();
    size_t cap = 16;
    while (cap < 4 * nb) cap <<= 1;
    if (edge_table.size() < cap) {
      edge_table.resize(cap);
      edge_stamp.assign(cap, 0);
    }
    const uint32_t gen = round;
    for (size_t b = 0; b < nb; ++b) {
      const int i = boundary[b].i;
      const Tet& old = olds[b];
      const int32_t nt = created[b];
      Tet T = old;
      T.v[i] = s;
      tets[old.n[i]].n[boundary[b].back] = nt;
      tets[nt] = T;
      for (int j = 0; j < 4; ++j) {
        if (j == i) continue;
        int32_t e[2], m = 0;
        for (int k = 0; k < 4; ++k)
          if (k != i && k != j) e[m++] = T.v[k];
        const uint64_t key = edge_key(e[0], e[1]);
        size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (cap - 1);
        for (;;) {
          if (edge_stamp[h] != gen) {
            edge_stamp[h] = gen;
            edge_table[h] = Slot{key, nt, j};
            break;
          }
          if (edge_table[h].key == key) {
            const Slot o2 = edge_table[h];
            tets[nt].n[j] = o2.tet;
            tets[o2.tet].n[o2.face] = nt;
            edge_table[h].key = ~0ull;  // each edge around s joins exactly two new tetrahedra
            break;
          }
          h = (h + 1) & (cap - 1);
        }
      }
    }
    if (!created.empty()) hint = created[0];
  }
};

// The Delaunay tetrahedralization of the n sites Pd [n][3] into T; dup_of[s]
// is the earlier site equal to s, or -1.
int triangulate(const double* Pd, int64_t n, Triangulation& T, std::vector<int32_t>& dup_of) {
  const double* P = Pd;
  // insertion order: a deterministic shuffle split into rounds of doubling
  // size (BRIO), each round sorted along a Morton curve of its bounding box
  std::vector<int32_t> order((size_t)n);
  for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
  uint64_t rng = 0x243F6A8885A308D3ull;
  for (int64_t i = n - 1; i > 0; --i) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    std::swap(order[i], order[rng % (uint64_t)(i + 1)]);
  }
  double lo[3] = {P[0], P[1], P[2]}, hi[3] = {P[0], P[1], P[2]};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], P[3 * i + k]);
      hi[k] = std::max(hi[k], P[3 * i + k]);
    }
  auto morton = [&](int32_t i) {
    uint64_t code = 0;
    uint32_t c[3];
    for (int k = 0; k < 3; ++k) {
      const double span = hi[k] > lo[k] ? hi[k] - lo[k] : 1.0;
      c[k] = (uint32_t)std::min(2097151.0, std::max(0.0, (P[3 * (size_t)i + k] - lo[k]) / span * 2097151.0));
    }
    for (int b = 20; b >= 0; --b)
      for (int k = 0; k < 3; ++k) code = code << 1 | ((c[k] >> b) & 1u);
    return code;
  };
  {
    std::vector<std::pair<uint64_t, int32_t>> keyed;
    int64_t start = 0, len = std::min<int64_t>(n, 1024);
    while (start < n) {
      const int64_t end = std::min<int64_t>(n, start + len);
      keyed.clear();
      for (int64_t i = start; i < end; ++i) keyed.push_back({morton(order[i]), order[i]});
      std::sort(keyed.begin(), keyed.end());
      for (int64_t i = start; i < end; ++i) order[i] = keyed[i - start].second;
      start = end;
      len *= 2;
    }
  }
  // initial tetrahedron: four affinely independent sites
  T.P = P;
  int32_t a = order[0], b = -1, c = -1, d = -1;
  size_t ib = 0, ic = 0;
  for (size_t k = 1; k < order.size() && b < 0; ++k) {
    const double* p = T.pt(order[k]);
    if (p[0] != T.pt(a)[0] || p[1] != T.pt(a)[1] || p[2] != T.pt(a)[2]) {
      b = order[k];
      ib = k;
    }
  }
  BV_REQUIRE(b >= 0, BV_ERR_INVALID, "3D Delaunay triangulation failed: all sites coincide");
  for (size_t k = 1; k < order.size() && c < 0; ++k) {
    if (k == ib) continue;
    // not collinear: the cross product of the two edges is not zero
    const double* pa = T.pt(a);
    const double* pb = T.pt(b);
    const double* pc = T.pt(order[k]);
    const double u[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    const double v[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
    const double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    if (std::fabs(cr[0]) + std::fabs(cr[1]) + std::fabs(cr[2]) > 1e-300) {
      c = order[k];
      ic = k;
    }
  }
  BV_REQUIRE(c >= 0, BV_ERR_INVALID, "3D Delaunay triangulation failed: all sites are collinear");
  int o = 0;
  for (size_t k = 1; k < order.size() && d < 0; ++k) {
    if (k == ib || k == ic) continue;
    o = orient(T.pt(a), T.pt(b), T.pt(c), T.pt(order[k]));
    if (o != 0) d = order[k];
  }
  BV_REQUIRE(d >= 0, BV_ERR_INVALID, "3D Delaunay triangulation failed: all sites are coplanar");
  if (o < 0) std::swap(a, b);  // positive orientation
  // the finite tetrahedron 0 and the four infinite ones across its faces:
  // infinite tetrahedron k is tetrahedron 0 with v[k] at infinity and two of
  // its finite vertices swapped, so it is positive with the infinite vertex
  // beyond face k
  T.tets.assign(5, Tet{});
  T.dead.assign(5, 0);
  T.stamp.assign(5, 0);
  T.tets[0] = Tet{{a, b, c, d}, {1, 2, 3, 4}};
  for (int k = 0; k < 4; ++k) {
    Tet t = T.tets[0];
    t.v[k] = kInf;
    std::swap(t.v[(k + 1) & 3], t.v[(k + 2) & 3]);
    for (int q = 0; q < 4; ++q) {
      if (t.v[q] == kInf) {
        t.n[q] = 0;  // across the hull face
        continue;
      }
      // the face opposite finite vertex t.v[q] holds kInf and the two finite
      // vertices other than T0.v[k] and t.v[q]: it is shared with the
      // infinite tetrahedron whose hull face lacks t.v[q]
      for (int m = 0; m < 4; ++m)
        if (T.tets[0].v[m] == t.v[q]) t.n[q] = 1 + m;
    }
    T.tets[1 + k] = t;
  }
  dup_of.assign((size_t)n, -1);
  for (size_t k = 0; k < order.size(); ++k) {
    const int32_t s = order[k];
    if (s == a || s == b || s == c || s == d) continue;
    int32_t dup = -1;
    T.insert(s, &dup);
    dup_of[s] = dup;
  }
  return BV_OK;
}

}  // namespace

namespace bvh {

// Delaunay edge graph of n sites (xyz [n][3] FP32, host).  Rows sorted and
// self-free; duplicate sites get empty rows.
int delaunay_csr(const float* xyz, int64_t n, std::vector<uint32_t>* offsets, std::vector<uint32_t>* neighbors) {
  BV_REQUIRE(n >= 4, BV_ERR_INVALID, "need at least 4 points for 3D Delaunay adjacency");
  BV_REQUIRE(n < (int64_t)0x7FFFFFFF, BV_ERR_INVALID, "too many sites");
  std::vector<double> P((size_t)n * 3);
  for (int64_t i = 0; i < 3 * n; ++i) {
    BV_REQUIRE(std::isfinite(xyz[i]), BV_ERR_INVALID, "non-finite site position");
    P[i] = (double)xyz[i];
  }
  Triangulation T;
  std::vector<int32_t> dup_of;
  const int rc = triangulate(P.data(), n, T, dup_of);
  if (rc != BV_OK) return rc;
  // edges of finite tetrahedra, each emitted from its lower endpoint, then
  // per-row sort and dedup, then both directions
  std::vector<uint32_t> cnt((size_t)n + 1, 0);
  static const int E[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (size_t t = 0; t < T.tets.size(); ++t) {
    if (T.dead[t] || T.has_inf(T.tets[t])) continue;
    for (auto& e : E) {
      const int32_t u = T.tets[t].v[e[0]], v = T.tets[t].v[e[1]];
      cnt[std::min(u, v) + 1] += 1;
    }
  }
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<uint32_t> half(cnt[n]);
  std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
  for (size_t t = 0; t < T.tets.size(); ++t) {
    if (T.dead[t] || T.has_inf(T.tets[t])) continue;
    for (auto& e : E) {
      const int32_t u = T.tets[t].v[e[0]], v = T.tets[t].v[e[1]];
      half[fill[std::min(u, v)]++] = (uint32_t)std::max(u, v);
    }
  }
  std::vector<uint32_t> deg((size_t)n, 0);
  std::vector<uint32_t> uniq_end((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    auto b0 = half.begin() + cnt[i], b1 = half.begin() + cnt[i + 1];
    std::sort(b0, b1);
    auto e1 = std::unique(b0, b1);
    uniq_end[i] = (uint32_t)(e1 - half.begin());
    deg[i] += (uint32_t)(e1 - b0);
    for (auto it = b0; it != e1; ++it) deg[*it] += 1;
  }
  offsets->assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < n; ++i) (*offsets)[i + 1] = (*offsets)[i] + deg[i];
  neighbors->assign((*offsets)[n], 0);
  std::vector<uint32_t> pos(offsets->begin(), offsets->end() - 1);
  // rows come out ascending: for row r, first every u < r (visited in
  // increasing u), then r's own upper neighbours in increasing order
  for (int64_t i = 0; i < n; ++i) {
    for (uint32_t k = cnt[i]; k < uniq_end[i]; ++k) {
      const uint32_t v = half[k];
      (*neighbors)[pos[v]++] = (uint32_t)i;
    }
  }
  for (int64_t i = 0; i < n; ++i)
    for (uint32_t k = cnt[i]; k < uniq_end[i]; ++k) (*neighbors)[pos[i]++] = half[k];
  // pos[i] now points past row i; rows mix "lower" entries (appended first)
  // and "upper" ones (appended last), both ascending and lower < i < upper
  return BV_OK;
}

}  // namespace bvh

extern "C" {

int bv_compute_delaunay(const float* xyz, int64_t n_sites, bv_adjacency** out) {
  BV_REQUIRE(out != nullptr && xyz != nullptr, BV_ERR_INVALID, "null argument");
  *out = nullptr;
  bv_adjacency* a = new bv_adjacency();
  a->n = n_sites;
  const int rc = bvh::delaunay_csr(xyz, n_sites, &a->host_offsets, &a->host_neighbors);
  if (rc != BV_OK) {
    delete a;
    return rc;
  }
  a->n_edges = a->host_neighbors.size();
  *out = a;
  return BV_OK;
}

}  // extern "C"
