// CPU oracle of the trainer's volumetric graph.  TEST INFRASTRUCTURE ONLY:
// only tests/ and bench.py's cpu_baseline leg load it, and only as the checker.
//
// A from-scratch restatement of build_volumetric_graph_with_options
// (blade-volume-train/src/diff_render.rs:943-1921) for plain RadFoam and
// PowerFoam paths without oriented surfaces, evaluated per pixel row:
//   density activation       :227-232 (ReLU, or softplus with beta), :1301-1302
//   dt from positions        :1327-1448 (bisector crossing with the bounded
//                            reciprocal d / (d^2 + 1e-12)), clamp to
//                            [0, MAX_PATH_DT] :1449-1457, normal gate blending
//                            the recorder's dt on terminal steps :1459-1473,
//                            :1557-1562
//   weighted linearization   :1474-1556 (dt_ref + tangent - tangent_ref with the
//                            recorded Jacobians of previous / current / next
//                            site geometry (xyz, softplus_100(log_radius)) minus
//                            the ray origin (xyz, 0)), clamped the same way
//   compositing              :1563-1583 (T = exp(-exclusive_cumsum(sigma dt)),
//                            alpha = 1 - exp(-sigma dt), w = T alpha)
//   colour                   appearance_rows :2162-2308 (relu(0.5 + sum coef
//                            basis)), pixel_sh :2311-2355, sh_basis :2062-2090
//                            with the trainer's constants :2001-2022
//   exposure, background     :1685-1712
//   losses                   L1 (mean absolute error per channel, summed),
//                            SmoothL1 (smooth_l1_loss :756-775, summed over
//                            channels and divided by 3) :1714-1735, opacity
//                            loss opacity_weight * mean((opacity - target)^2)
//                            :1763-1775; L2 (mean squared error per channel,
//                            summed) is this build's extension for the L2
//                            training workload.
// The gradient is the exact reverse mode of that forward (ReLU derivative 0
// at 0, sign(0) = 0 for L1).  Templated on the scalar: float is the parity
// oracle, double the finite-difference twin.  Compiled with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kDtPositions = 0;   // RadFoam: dt from the bisector planes
constexpr int kDtRecorded = 1;    // weighted path, frozen geometry: the recorder's dt
constexpr int kDtLinearized = 2;  // weighted path with recorded Jacobians
constexpr int kLossL1 = 0, kLossSmoothL1 = 1, kLossL2 = 2;

// diff_render.rs:2001-2022, f32 literals (the double twin promotes them).
const float kC0 = 0.2820948f;
const float kC1 = 0.4886025f;
const float kC2[5] = {1.0925484f, -1.0925484f, 0.31539157f, -1.0925484f, 0.5462742f};
const float kC3[7] = {-0.5900436f, 2.8906114f, -0.4570458f, 0.37317634f, -0.4570458f, 1.4453057f, -0.5900436f};

// sh_basis, diff_render.rs:2062-2090 (same products, same order).
template <class T> void sh_basis(const T d[3], int K, T out[16]) {
  const T x = d[0], y = d[1], z = d[2];
  out[0] = T(kC0);
  if (K > 1) {
    out[1] = -T(kC1) * y;
    out[2] = T(kC1) * z;
    out[3] = -T(kC1) * x;
  }
  if (K > 4) {
    const T xx = x * x, yy = y * y, zz = z * z;
    out[4] = T(kC2[0]) * x * y;
    out[5] = T(kC2[1]) * y * z;
    out[6] = T(kC2[2]) * (T(3) * zz - T(1));
    out[7] = T(kC2[3]) * x * z;
    out[8] = T(kC2[4]) * (xx - yy);
  }
  if (K > 9) {
    const T xx = x * x, yy = y * y, zz = z * z;
    out[9] = T(kC3[0]) * y * (T(3) * xx - yy);
    out[10] = T(kC3[1]) * x * y * z;
    out[11] = T(kC3[2]) * y * (T(5) * zz - T(1));
    out[12] = T(kC3[3]) * z * (T(5) * zz - T(3));
    out[13] = T(kC3[4]) * x * (T(5) * zz - T(1));
    out[14] = T(kC3[5]) * z * (xx - yy);
    out[15] = T(kC3[6]) * x * (xx - T(3) * yy);
  }
}

template <class T> struct Params {
  int64_t n;
  int sh_degree;
  int n_views;
  const T* log_density;  // [n]
  const T* positions;    // [n][3]
  const T* sh;           // [n][3K], RGB interleaved per component
  const T* log_radii;    // [n] or null
  const T* exposure;     // [n_views][3] or null (= 1)
};

template <class T> struct Batch {
  int64_t P;
  int L;
  const uint32_t* cells;
  const uint32_t* next;
  const T* dt_rec;
  const T* mask;
  const uint32_t* status;  // [P] or null: rows hold status & 0x7fffffff steps, the rest is padding
  const uint32_t* prev;
  const T* ref_tangent;
  const T* jac_prev;  // [P*L][4]
  const T* jac_cur;
  const T* jac_next;
  const T* origin;   // [P][3]
  const T* dir;      // [P][3]
  const T* basis;    // [P][K-1] or null
  const uint32_t* view;
  const T* labels;   // [P][3]
  const T* target_alpha;
};

template <class T> struct Options {
  T softplus_beta;
  T background[3];
  int loss;
  T opacity_weight;
  int dt_mode;
};

template <class T> struct Grads {
  T* d_log_density;
  T* d_positions;
  T* d_sh;
  T* d_log_radii;
  T* d_exposure;
};

template <class T> inline T relu(T x) { return x > T(0) ? x : T(0); }
template <class T> inline T step01(T x) { return x > T(0) ? T(1) : T(0); }

// Stable softplus with slope beta: log(1 + exp(beta x)) / beta.
template <class T> inline T softplus(T x, T beta) {
  const T bx = beta * x;
  return (relu(bx) + std::log1p(std::exp(-std::fabs(bx)))) / beta;
}
template <class T> inline T sigmoid(T x) { return T(1) / (T(1) + std::exp(-x)); }

template <class T> inline T activate(T x, T beta) { return beta > T(0) ? softplus(x, beta) : relu(x); }
template <class T> inline T activate_grad(T x, T beta) { return beta > T(0) ? sigmoid(beta * x) : step01(x); }

const float kMaxPathDt = 50.0f;     // diff_render.rs:2510
const float kRadiusBeta = 100.0f;   // RADIUS_SOFTPLUS_BETA, :223

// One step's recomputed forward state.
template <class T> struct Step {
  bool in;
  uint32_t cell, next, prev;
  T mask;
  // dt from positions
  T pc[3], pn[3], n[3], mo[3];
  T num, den, nn, t, gate, dt_raw, dt_pos, dt_clamped;
  // weighted
  T lin;
  T dt_sel, dt;
  T x, dens, raw;
  T pre[3], col[3];
};

template <class T> T dot3(const T a[3], const T b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

template <class T>
void step_forward(const Params<T>& prm, const Batch<T>& b, const Options<T>& opt, int64_t p, int k, const T* bas,
                  int K, T t_prev, Step<T>* s) {
  const size_t i = (size_t)p * b.L + k;
  s->cell = b.cells[i];
  s->next = b.next[i];
  s->prev = b.prev ? b.prev[i] : s->cell;
  s->mask = b.mask[i];
  const T* o = b.origin + 3 * p;
  const T* d = b.dir + 3 * p;
  if (opt.dt_mode == kDtPositions) {
    // :1327-1448
    for (int c = 0; c < 3; ++c) {
      s->pc[c] = prm.positions[3 * (size_t)s->cell + c];
      s->pn[c] = prm.positions[3 * (size_t)s->next + c];
      const T mid = (s->pc[c] + s->pn[c]) * T(0.5);
      s->n[c] = s->pn[c] + (-s->pc[c]);
      s->mo[c] = mid + (-o[c]);
    }
    s->nn = dot3(s->n, s->n);
    s->num = dot3(s->mo, s->n);
    s->den = dot3(s->n, d);
    s->t = (s->num * s->den) / (s->den * s->den + T(1e-12f));
    s->dt_raw = s->t + (-t_prev);
    s->dt_pos = relu(s->dt_raw);
    const T over = relu(T(kMaxPathDt) + (-s->dt_pos));
    s->dt_clamped = T(kMaxPathDt) + (-over);
    s->gate = s->nn / (s->nn + T(1e-6f));
    s->dt_sel = s->dt_clamped * s->gate + b.dt_rec[i] * (T(1) + (-s->gate));
  } else if (opt.dt_mode == kDtLinearized) {
    // :1474-1556
    const uint32_t roles[3] = {s->prev, s->cell, s->next};
    const T* jac[3] = {b.jac_prev + 4 * i, b.jac_cur + 4 * i, b.jac_next + 4 * i};
    T term[3];
    for (int r = 0; r < 3; ++r) {
      T g[4];
      for (int c = 0; c < 3; ++c) g[c] = prm.positions[3 * (size_t)roles[r] + c] + (-o[c]);
      g[3] = softplus(prm.log_radii[roles[r]], T(kRadiusBeta)) + (-T(0));
      term[r] = ((g[0] * jac[r][0] + g[1] * jac[r][1]) + g[2] * jac[r][2]) + g[3] * jac[r][3];
    }
    const T geometry = (term[0] + term[1]) + term[2];
    const T delta = geometry + (-b.ref_tangent[i]);
    s->lin = b.dt_rec[i] + delta;
    const T pos = relu(s->lin);
    const T within = relu(T(kMaxPathDt) + (-pos));
    s->dt_sel = T(kMaxPathDt) + (-within);
  } else {
    s->dt_sel = b.dt_rec[i];
  }
  s->dt = s->dt_sel * s->mask;
  s->x = prm.log_density[s->cell];
  s->dens = activate(s->x, opt.softplus_beta);
  s->raw = s->dens * s->dt;
  // appearance_rows: per channel sum_j coef_j basis_j, + 0.5, relu
  const T* row = prm.sh + (size_t)s->cell * 3 * K;
  for (int c = 0; c < 3; ++c) {
    T acc = T(0);
    for (int j = 0; j < K; ++j) acc += row[3 * j + c] * bas[j];
    s->pre[c] = acc + T(0.5);
    s->col[c] = relu(s->pre[c]);
  }
}

template <class T> inline void add3(T* g, uint32_t cell, const T v[3]) {
  if (!g) return;
  for (int c = 0; c < 3; ++c) g[3 * (size_t)cell + c] += v[c];
}

// Forward and reverse pass of one pixel row.  Returns the row's loss terms;
// gradients are added into `gr` (serial per row: the caller runs rows in order).
template <class T>
T graph_row(const Params<T>& prm, const Batch<T>& b, const Options<T>& opt, int64_t p, const Grads<T>& gr,
            T* pixel_out, T* opacity_out, T* dt_out) {
  const int K = (prm.sh_degree + 1) * (prm.sh_degree + 1);
  T bas[16];
  if (b.basis) {
    bas[0] = T(kC0);
    for (int j = 1; j < K; ++j) bas[j] = b.basis[(size_t)p * (K - 1) + (j - 1)];
  } else {
    sh_basis(b.dir + 3 * p, K, bas);
  }
  int Lr = b.L;
  if (b.status) {
    const int cnt = (int)(b.status[p] & 0x7fffffffu);
    Lr = cnt < b.L ? cnt : b.L;
  }
  const uint32_t view = b.view ? b.view[p] : 0u;
  T E[3] = {T(1), T(1), T(1)};
  if (prm.exposure)
    for (int c = 0; c < 3; ++c) E[c] = prm.exposure[3 * (size_t)view + c];
  // forward
  std::vector<Step<T>> st((size_t)(Lr > 0 ? Lr : 0));
  std::vector<T> cum(st.size()), Tk(st.size()), alpha(st.size()), w(st.size());
  T t_prev = T(0), run = T(0), P[3] = {T(0), T(0), T(0)}, O = T(0);
  for (int k = 0; k < Lr; ++k) {
    step_forward(prm, b, opt, p, k, bas, K, t_prev, &st[k]);
    t_prev = st[k].t;
    cum[k] = run;
    run += st[k].raw;
    Tk[k] = std::exp(-cum[k]);
    alpha[k] = T(1) + (-std::exp(-st[k].raw));
    w[k] = Tk[k] * alpha[k];
    for (int c = 0; c < 3; ++c) P[c] += w[k] * st[k].col[c];
    O += w[k];
    if (dt_out) dt_out[(size_t)p * b.L + k] = st[k].dt;
  }
  if (dt_out)
    for (int k = Lr; k < b.L; ++k) dt_out[(size_t)p * b.L + k] = T(0);
  // pixel, loss, dL/dpixel
  const T inv_p = T(1) / T(b.P);
  T pix[3], g[3], loss = T(0);
  for (int c = 0; c < 3; ++c) {
    pix[c] = P[c] * E[c];
    if (opt.background[c] != T(0)) pix[c] = pix[c] + (T(1) + (-O)) * opt.background[c];
    const T diff = pix[c] + (-b.labels[3 * p + c]);
    if (opt.loss == kLossL1) {
      loss += std::fabs(diff) * inv_p;
      g[c] = (diff > T(0) ? T(1) : (diff < T(0) ? T(-1) : T(0))) * inv_p;
    } else if (opt.loss == kLossSmoothL1) {
      const T a = std::fabs(diff);
      const T above = relu(a + T(-1));
      const T cl = a + (-above);
      loss += (cl * cl * T(0.5) + above) * inv_p / T(3);
      const T da = a > T(1) ? T(1) : a;  // d/da (0.5 clamp^2 + relu(a - 1))
      g[c] = (diff > T(0) ? da : (diff < T(0) ? -da : T(0))) * inv_p / T(3);
    } else {
      loss += diff * diff * inv_p;
      g[c] = T(2) * diff * inv_p;
    }
    pixel_out[3 * p + c] = pix[c];
  }
  T dop = T(0);
  if (opt.opacity_weight > T(0)) {
    const T e = O + (-b.target_alpha[p]);
    loss += opt.opacity_weight * e * e * inv_p;
    dop = opt.opacity_weight * T(2) * e * inv_p;
  }
  if (opacity_out) opacity_out[p] = O;
  if (gr.d_exposure)
    for (int c = 0; c < 3; ++c) gr.d_exposure[3 * (size_t)view + c] += g[c] * P[c];
  // reverse: dL/dw_k = sum_c g_c (E_c c_kc - bg_c) + dop; dT_k/draw_j = -T_k for j < k
  std::vector<T> gw(st.size());
  for (int k = 0; k < Lr; ++k) {
    T v = dop;
    for (int c = 0; c < 3; ++c) v += g[c] * (E[c] * st[k].col[c] - opt.background[c]);
    gw[k] = v;
  }
  T suffix = T(0);  // sum_{j > k} gw_j w_j
  std::vector<T> d_dt_raw(st.size(), T(0));
  for (int k = Lr - 1; k >= 0; --k) {
    const Step<T>& s = st[k];
    const T T_next = Tk[k] * (T(1) - alpha[k]);
    const T d_raw = gw[k] * T_next - suffix;
    suffix += gw[k] * w[k];
    // density
    const T d_dens = d_raw * s.dt;
    if (gr.d_log_density) gr.d_log_density[s.cell] += d_dens * activate_grad(s.x, opt.softplus_beta);
    // colour
    if (gr.d_sh) {
      for (int c = 0; c < 3; ++c) {
        const T dc = g[c] * E[c] * w[k] * step01(s.pre[c]);
        for (int j = 0; j < K; ++j) gr.d_sh[(size_t)s.cell * 3 * K + 3 * j + c] += dc * bas[j];
      }
    }
    const T d_sel = d_raw * s.dens * s.mask;
    if (opt.dt_mode == kDtPositions) {
      d_dt_raw[k] = d_sel * s.gate * step01(s.dt_raw) * step01(T(kMaxPathDt) + (-s.dt_pos));
      const T d_gate = d_sel * (s.dt_clamped + (-b.dt_rec[(size_t)p * b.L + k]));
      const T bb = s.nn + T(1e-6f);
      const T d_nn = d_gate * T(1e-6f) / (bb * bb);
      // t_k feeds dt_raw_k (+) and dt_raw_{k+1} (-)
      const T d_t = d_dt_raw[k] - (k + 1 < Lr ? d_dt_raw[k + 1] : T(0));
      const T q = s.den * s.den + T(1e-12f);
      const T d_num = d_t * s.den / q;
      const T d_den = d_t * s.num * (T(1e-12f) - s.den * s.den) / (q * q);
      const T* d = b.dir + 3 * p;
      T gc[3], gn[3];
      for (int c = 0; c < 3; ++c) {
        const T dn = d_num * s.mo[c] + d_den * d[c] + T(2) * d_nn * s.n[c];
        const T dmo = d_num * s.n[c];
        gc[c] = T(0.5) * dmo - dn;
        gn[c] = T(0.5) * dmo + dn;
      }
      add3(gr.d_positions, s.cell, gc);
      add3(gr.d_positions, s.next, gn);
    } else if (opt.dt_mode == kDtLinearized) {
      const T d_lin = d_sel * step01(s.lin) * step01(T(kMaxPathDt) + (-relu(s.lin)));
      const size_t i = (size_t)p * b.L + k;
      const uint32_t roles[3] = {s.prev, s.cell, s.next};
      const T* jac[3] = {b.jac_prev + 4 * i, b.jac_cur + 4 * i, b.jac_next + 4 * i};
      for (int r = 0; r < 3; ++r) {
        const T v[3] = {d_lin * jac[r][0], d_lin * jac[r][1], d_lin * jac[r][2]};
        add3(gr.d_positions, roles[r], v);
        if (gr.d_log_radii)
          gr.d_log_radii[roles[r]] +=
              d_lin * jac[r][3] * sigmoid(T(kRadiusBeta) * prm.log_radii[roles[r]]);
      }
    }
  }
  return loss;
}

template <class T>
T graph_step(const Params<T>& prm, const Batch<T>& b, const Options<T>& opt, const Grads<T>& gr, T* pixels,
             T* opacity, T* dt_out) {
  T loss = T(0);
  for (int64_t p = 0; p < b.P; ++p) loss += graph_row(prm, b, opt, p, gr, pixels, opacity, dt_out);
  return loss;
}

}  // namespace

extern "C" {

// Flat C views of the structs above; one set per scalar type.
#define BVO_GRAPH_STRUCTS(T, S)                                                                             \
  typedef struct {                                                                                          \
    int64_t n;                                                                                              \
    int32_t sh_degree, n_views;                                                                             \
    const T *log_density, *positions, *sh, *log_radii, *exposure;                                           \
  } bvo_graph_params_##S;                                                                                   \
  typedef struct {                                                                                          \
    int64_t P;                                                                                              \
    int32_t L, pad;                                                                                         \
    const uint32_t *cells, *next;                                                                           \
    const T *dt_rec, *mask;                                                                                 \
    const uint32_t *status, *prev;                                                                          \
    const T *ref_tangent, *jac_prev, *jac_cur, *jac_next, *origin, *dir, *basis;                            \
    const uint32_t* view;                                                                                   \
    const T *labels, *target_alpha;                                                                         \
  } bvo_graph_batch_##S;                                                                                    \
  typedef struct {                                                                                          \
    T softplus_beta;                                                                                        \
    T background[3];                                                                                        \
    int32_t loss;                                                                                           \
    T opacity_weight;                                                                                       \
    int32_t dt_mode;                                                                                        \
  } bvo_graph_options_##S;                                                                                  \
  typedef struct {                                                                                          \
    T *d_log_density, *d_positions, *d_sh, *d_log_radii, *d_exposure;                                       \
  } bvo_graph_grads_##S;

BVO_GRAPH_STRUCTS(float, f32)
BVO_GRAPH_STRUCTS(double, f64)

#define BVO_GRAPH_ENTRY(T, S)                                                                                  \
  T bvo_graph_step_##S(const bvo_graph_params_##S* pp, const bvo_graph_batch_##S* bp,                         \
                       const bvo_graph_options_##S* op, const bvo_graph_grads_##S* gp, T* pixels, T* opacity,  \
                       T* dt_out) {                                                                            \
    Params<T> prm{pp->n, pp->sh_degree, pp->n_views, pp->log_density, pp->positions, pp->sh, pp->log_radii,    \
                  pp->exposure};                                                                               \
    Batch<T> b{bp->P,       bp->L,          bp->cells,    bp->next,    bp->dt_rec, bp->mask,                   \
               bp->status,  bp->prev,       bp->ref_tangent, bp->jac_prev, bp->jac_cur, bp->jac_next,          \
               bp->origin,  bp->dir,        bp->basis,    bp->view,    bp->labels, bp->target_alpha};          \
    Options<T> opt{op->softplus_beta, {op->background[0], op->background[1], op->background[2]}, op->loss,     \
                   op->opacity_weight, op->dt_mode};                                                           \
    Grads<T> gr{gp->d_log_density, gp->d_positions, gp->d_sh, gp->d_log_radii, gp->d_exposure};                \
    return graph_step(prm, b, opt, gr, pixels, opacity, dt_out);                                               \
  }

BVO_GRAPH_ENTRY(float, f32)
BVO_GRAPH_ENTRY(double, f64)

// The trainer's SH basis at one direction (sh_basis, diff_render.rs:2062-2090).
void bvo_graph_sh_basis_f32(const float dir[3], int32_t K, float out[16]) { sh_basis(dir, K, out); }

}  // extern "C"
