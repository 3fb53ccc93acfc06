/* blade-volume on MI355X: C ABI of the RadFoam / PowerFoam cell walk.
 *
 * This header is the drop-in boundary for the reference's GPU hot path.  Each
 * entry point names the reference interface it replaces (paths relative to the
 * reference repository root).  Conventions:
 *   - every function returns an int status (BV_OK on success) and never
 *     unwinds; bv_last_error() gives a thread-local description;
 *   - handles are opaque; the library owns the scene's device memory, the
 *     caller owns every other buffer;
 *   - buffers passed to trace / record / adjoint calls are DEVICE pointers
 *     (HBM-resident, e.g. from hipMalloc or a torch tensor); scene creation
 *     takes HOST arrays and uploads them;
 *   - `stream` is a hipStream_t (NULL = the default stream); calls are
 *     asynchronous on it unless stated otherwise;
 *   - distinct scene handles may be used from distinct threads.
 */
#ifndef BV_H
#define BV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BV_OK 0
#define BV_ERR_INVALID 1      /* bad argument or model (reference: assert!/validate() panics) */
#define BV_ERR_UNSUPPORTED 2  /* oriented / surface-detail / spherical-Voronoi flags, or a mode this build lacks */
#define BV_ERR_HIP 3          /* a HIP runtime call failed */
#define BV_ERR_OOM 4          /* device allocation failed */
#define BV_ERR_CAPACITY 5     /* a caller-supplied capacity is too small */

/* Scene flags that the reference supports and this build rejects (lib.rs:187-239). */
#define BV_SCENE_ORIENTED 0x1u
#define BV_SCENE_SURFACE_DETAIL 0x2u
#define BV_SCENE_SPHERICAL_VORONOI 0x4u
#define BV_SCENE_SURFACE_COLOR 0x8u

/* Walk kernels.  PLAIN is the one-ray-per-thread port of the reference WGSL
 * walk (radfoam_trace.wgsl:104-219): AoS float4 points, one gather per
 * neighbour.  FAST is the MI355X walk over the per-edge plane layout.  Both
 * produce identical cell sequences. */
#define BV_KERNEL_PLAIN 0
#define BV_KERNEL_FAST 1

/* Colour evaluation, trace.rs:38-48 (EvalMode), plus point-light relighting. */
#define BV_EVAL_OPACITY 0
#define BV_EVAL_CONSTANT 1
#define BV_EVAL_SH 2
#define BV_EVAL_RELIGHT 3   /* per-cell material lit by 1..8 point lights (bv_lighting) */
#define BV_MAX_LIGHTS 8
#define BV_MAX_OBJECTS 8

#define BV_NO_CELL 0xFFFFFFFFu
#define BV_PATH_TRUNCATED 0x80000000u

typedef struct bv_scene bv_scene;

/* PointCloudModel (lib.rs:173-239) as SoA host arrays. */
typedef struct {
  int64_t n_sites;
  const float* xyz;          /* [n][3] */
  const float* radius;       /* [n] or NULL: NULL = plain RadFoam (unbounded Voronoi cells) */
  const float* density;      /* [n] */
  const float* sh;           /* [n][3K], K = (min(sh_degree,3)+1)^2, RGB interleaved per component */
  int32_t sh_degree;
  const uint32_t* offsets;   /* [n+1] CSR */
  const uint32_t* neighbors; /* [offsets[n]] */
  uint32_t flags;            /* BV_SCENE_*; any set flag -> BV_ERR_UNSUPPORTED */
  const float* material;     /* [n][4] albedo rgb + GGX roughness (relight.rs:77-85), or NULL */
} bv_scene_desc;

typedef struct {
  int64_t n_sites;
  int64_t n_edges;
  int32_t sh_degree;
  int32_t bounded;           /* 1 when radii are present (PowerFoam) */
  uint32_t max_degree;
  uint32_t pad;
  int64_t device_bytes;      /* HBM held by the scene */
} bv_scene_info;

/* TraceSettings (trace.rs:50-69) / RadFoamTraceSettings (gpu/radfoam_trace.rs:4-23). */
typedef struct {
  float weight_threshold;    /* stop when transmittance <= this */
  uint32_t max_steps;
  float depth;               /* far plane (camera depth) */
  uint32_t eval_mode;        /* BV_EVAL_* */
  float const_rgb[3];        /* colour for BV_EVAL_CONSTANT */
  uint32_t pad;
} bv_trace_settings;

/* A batch of rays in device memory. */
typedef struct {
  const float* origins;        /* [count][3] */
  const float* directions;     /* [count][3], need not be unit length (trace.rs:1704-1706 normalises) */
  const uint32_t* start_cells; /* [count] or NULL to use start_cell for every ray; in the walk entry points
                                  (trace, record, relight, adjoint, training) an entry >= n_sites, BV_NO_CELL
                                  included, makes that ray empty: zero rgba, steps 0, t_end 0, status 0, no
                                  gradient, no cell read */
  uint32_t start_cell;
  uint32_t pad;
  int64_t count;               /* at most 2^32 - 1 rays per call */
} bv_rays;

/* TraceResult (trace.rs:71-103) as device SoA outputs; any pointer may be NULL. */
typedef struct {
  float* rgba;               /* [count][4] rgb + alpha (1 - transmittance), FP32 */
  uint32_t* steps;           /* [count] iterations executed (trace.rs semantics) */
  uint32_t* last_point;      /* [count] */
  float* t_end;              /* [count] */
  float* depth_mode;         /* [count] */
  float* peak_weight;        /* [count] */
  uint32_t* peak_point;      /* [count], BV_NO_CELL when no positive-density segment */
  uint32_t* stats;           /* [count][2] instrumentation: algorithmic bytes of the walk
                                (per visited cell 8 + 16 deg (+4 deg radii) + 4 + 12K if integrated,
                                plus 16 for the rgba store) and the number of integrated segments */
  uint32_t* path;            /* [count][path_len] the visited-cell sequence: the current cell of each of the
                                walk's first path_len iterations (the one that stops included); slots past
                                steps are not written.  Cell walks only (bv_trace_rays / _camera). */
  uint32_t path_len;
  uint32_t pad;
} bv_trace_outputs;

/* CameraParams (camera.rs:9-18), 48 bytes. */
typedef struct {
  float position[3];
  float depth;
  float orientation[4];      /* quaternion x, y, z, w */
  float fov[2];              /* (fov_x, fov_y) radians */
  float principal[2];        /* NDC; zero = image centre */
} bv_camera;

/* PointLight (relight.rs:176-181): a finite light with angular falloff
 * max(dot(direction, away), 0)^exponent (exponent 0 = isotropic). */
typedef struct {
  float position[3];
  float exponent;
  float direction[3];
  uint32_t visibility;       /* 1: this light casts a visibility ray (see bv_lighting) */
  float intensity[3];
  float pad1;
} bv_point_light;

/* Lighting of a relit render.  Each integrated segment is shaded with
 * shade_point_light (relight.rs:295-324) at its entry point x, normal x - site,
 * view -d and the cell's material; for a light with visibility on (its own
 * flag, or lighting->visibility for every light), its contribution is scaled
 * by the transmittance of an opacity walk from the ray's depth point toward
 * the light (the reference's visibility ray, relight.wgsl:212-229). */
typedef struct {
  uint32_t count;            /* 1..BV_MAX_LIGHTS */
  uint32_t visibility;       /* 1: every light casts a visibility ray; 0: only lights whose own flag is set */
  float specular_f0[3];      /* Material::specular_f0 for every cell; 0 = diffuse only */
  float pad;
  bv_point_light lights[BV_MAX_LIGHTS];
} bv_lighting;

/* RecordPathsArgs subset (gpu/path_record.rs:142-161). */
typedef struct {
  uint32_t max_steps;        /* L: row length */
  float depth;
  float max_path_dt;         /* dt clamp (MAX_PATH_DT = 50 in the trainer) */
  uint32_t pad;
} bv_record_settings;

/* Gradient tables in device memory; the adjoint ADDS into them. */
typedef struct {
  float* d_density;          /* [n] */
  float* d_sh;               /* [n][3K] */
  float* d_xyz;              /* [n][3] */
} bv_grads;

const char* bv_last_error(void);
int bv_version(void);

/* Device info: number of visible GPUs; sets *cu_count for device 0. */
int bv_device_info(int* device_count, int* cu_count);

/* Replaces RadFoamGpuCloud::new (gpu/radfoam.rs:119-125, new_inner :139-533):
 * validates the model (lib.rs:246-483 subset), uploads it and builds the
 * per-edge plane layout.  Synchronous. */
int bv_scene_create(const bv_scene_desc* desc, bv_scene** out_scene);
/* Replaces RadFoamGpuCloud::deinit (gpu/radfoam.rs:535-543). */
int bv_scene_destroy(bv_scene* scene);
int bv_scene_get_info(const bv_scene* scene, bv_scene_info* out_info);

/* Refresh trainable parameters from DEVICE buffers (any pointer may be NULL)
 * and rebuild the derived layouts.  The trainer-side counterpart of
 * RadFoamGpuCloud::update_surface_geometry (gpu/radfoam.rs:559-637): the
 * adjacency is kept, only values change. */
int bv_scene_update(bv_scene* scene, const float* xyz, const float* density, const float* sh, void* stream);

/* Replaces PointIndex::containing_point (gpu/radfoam.rs:77-112) and
 * PointCloudModel::containing_cell (lib.rs:526-541): exact (power-)nearest
 * site, lowest index on ties.  position is a host array; synchronous on
 * `stream` (returns once the cell is known).  The reduction slot is per
 * stream, so queries of one scene on different streams do not interfere. */
int bv_containing_cell(bv_scene* scene, const float position[3], uint32_t* out_cell, void* stream);

/* Adjacency (CSR) built on the device; owned by the library until destroyed. */
typedef struct bv_adjacency bv_adjacency;

/* Replaces adjacency::compute_cech (adjacency.rs:406-425, the rows of
 * compute_cech_with_workers :427-540, sorted by build_symmetric_csr :55-85)
 * with the default AdjacencyConfig (unbounded, exact rows): the Čech graph of
 * weighted sites.  Edge {i, j}, i != j, whenever the FP32 squared distance
 * ((dx dx + dy dy) + dz dz) <= (r_i + r_j)^2, negative and NaN radii taken
 * as 0; every row ascending.  xyz [n][3] and radius [n] are HOST arrays; the
 * call is synchronous on `stream`.  Fails with BV_ERR_UNSUPPORTED beyond
 * 2^31 - 1 directed edges. */
int bv_compute_cech(const float* xyz, const float* radius, int64_t n_sites, void* stream, bv_adjacency** out);
/* Replaces adjacency::try_compute_adjacency (adjacency.rs:142-197; the
 * simple_delaunay_lib backend, or Qhull :564-634) with the default
 * AdjacencyConfig: the edge graph of the sites' 3-D Delaunay tetrahedralization
 * (edges of finite tetrahedra), every row ascending and self-free.  Exact
 * predicates; cospherical sites are resolved by a symbolic perturbation and
 * a site equal to an earlier one gets an empty row.  xyz [n][3] is a HOST
 * array; the builder runs on the host (no GPU needed) and the result is read
 * with bv_adjacency_read.  Fails (BV_ERR_INVALID) below 4 sites, on
 * non-finite positions and when all sites are coplanar
 * (AdjacencyError::TooFewPoints / TriangulationFailed). */
int bv_compute_delaunay(const float* xyz, int64_t n_sites, bv_adjacency** out);
/* Site count and directed edge count (the length of the neighbour array). */
int bv_adjacency_get_info(const bv_adjacency* adjacency, int64_t* n_sites, uint64_t* n_edges);
/* Copies the CSR into HOST arrays offsets [n + 1] and neighbors [n_edges]
 * (either may be NULL); synchronous.  The Adjacency { neighbors, offsets }
 * of lib.rs:98-103. */
int bv_adjacency_read(const bv_adjacency* adjacency, uint32_t* offsets, uint32_t* neighbors);
int bv_adjacency_destroy(bv_adjacency* adjacency);

/* Replaces trace::trace_one_ray over a batch (trace.rs:1680-1820) — the walk
 * that RadFoamGpuTracer dispatches (radfoam_trace.wgsl:104-219). */
int bv_trace_rays(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                  const bv_trace_outputs* outputs, int kernel, void* stream);

/* One RadFoam / PowerFoam object of a multi-object scene: ObjectBounds and
 * GpuTransform (scene.rs:62-100).  world_to_object is the inverse of the
 * object's Transform::to_matrix as a row-major 3x4 affine map,
 * p_obj[r] = m[4r] x + m[4r+1] y + m[4r+2] z + m[4r+3]; center / radius are the
 * world bounding sphere; start_cell is the object's cell containing the camera
 * (Scene::update_camera, scene.rs:395-413, via bv_containing_cell at the
 * camera's object-space position). */
typedef struct {
  bv_scene* scene;
  float world_to_object[12];
  float center[3];
  float radius;
  uint32_t start_cell;
  uint32_t pad;
} bv_scene_object;

/* Replaces trace_scene (scene_trace_core.wgsl:406-465) for RadFoam objects:
 * the ray (direction normalised) visits the objects whose bounding sphere it
 * hits in (t_near, index) order (find_next_object_hit, :55-74; t_near clamped
 * to 0, ray_sphere_intersect :14-41); each is walked in object space with the
 * transformed, unnormalised direction, so t stays the world distance, from its
 * start cell, integrating only inside [t_near, t_far]
 * (scene_trace_object, :378-402); results composite front to back
 * (rgb += T rgb_k, T *= 1 - alpha_k) until T < weight_threshold.  out_rgba is
 * [count][4] (rgb, 1 - T).  All objects share one SH degree and are all plain
 * RadFoam or all PowerFoam (else BV_ERR_UNSUPPORTED); eval_mode BV_EVAL_SH or
 * BV_EVAL_CONSTANT; settings->depth is unused (each object ends at t_far).
 * rays->start_cell(s) are unused. */
int bv_trace_objects(const bv_scene_object* objects, uint32_t n_objects, const bv_rays* rays,
                     const bv_trace_settings* settings, float* out_rgba, void* stream);

/* Replaces RadFoamGpuTracer::dispatch (gpu/radfoam_trace.rs:106-135, trace_main
 * radfoam.wgsl:38-56): one ray per pixel centre of a width x height image,
 * outputs indexed y * width + x.  start_cell = BV_NO_CELL computes the
 * camera's containing cell first (as the reference does on the host). */
int bv_trace_camera(bv_scene* scene, const bv_camera* camera, uint32_t width, uint32_t height,
                    uint32_t start_cell, const bv_trace_settings* settings, const bv_trace_outputs* outputs,
                    int kernel, void* stream);

/* Pixel-centre camera rays into device buffers (radfoam.wgsl:45-52 with
 * common.wgsl:19-23 qrot); the exact rays bv_trace_camera walks.  The
 * half-angle tangent is (float)tan((double)(0.5f * fov)), one rounding. */
int bv_camera_rays(const bv_camera* camera, uint32_t width, uint32_t height, float* origins,
                   float* directions, void* stream);

/* Replaces PathRecorder::dispatch for plain RadFoam (gpu/path_record.rs:433-495,
 * record_paths radfoam_record_paths.wgsl:1220-1355): [P, L] rows of cell,
 * next cell, dt (clamped to max_path_dt), mask = 1; status[p] = count |
 * BV_PATH_TRUNCATED.  Slots past count are not written (the reference leaves
 * them at their pre-dispatch value).  mask may be NULL. */
int bv_record_paths(bv_scene* scene, const bv_rays* rays, const bv_record_settings* settings,
                    uint32_t* cells, uint32_t* next_cells, float* dts, float* mask, uint32_t* status,
                    void* stream);

/* Replaces PathCompactor::dispatch (gpu/path_compact.rs:63-89, compact_paths in
 * path_compact.wgsl:24-50): appends every path's live prefix (status[p] &
 * 0x7fffffff records of its max_steps-wide row in `cells`, as bv_record_paths
 * writes them) to the dense stream at *count, one record per step:
 * dense_slots = p * max_steps + step, compact_cells = cells[that slot],
 * pixel_indices = p, active = 1.  *count (device) is advanced by the records
 * appended; the caller zeroes it (and `active`) before a fresh stream, as the
 * reference requires.  Paths are appended in path order, one of the orders the
 * reference's per-path atomic can produce, identical on every run.  Records at
 * or past `capacity` are not written (*count still counts them).
 * num_pixels * max_steps must fit 32 bits; asynchronous on `stream`. */
int bv_compact_paths(const uint32_t* path_status, const uint32_t* cells, uint32_t num_pixels, uint32_t max_steps,
                     uint32_t* count, uint32_t* dense_slots, uint32_t* compact_cells, uint32_t* pixel_indices,
                     float* active, uint64_t capacity, void* stream);

/* Adjoint of bv_trace_rays in BV_EVAL_SH mode for plain RadFoam: the exact
 * gradient of trace_one_ray (trace.rs:1680-1820) with respect to density, SH
 * coefficients and site positions, by a replay of the same walk (transmittance
 * early-out, raw density, unclamped dt, trace.rs's SH constants; position
 * terms through the bisector-plane crossings, face_intersection_jacobians
 * trace.rs:1077-1097).  It is NOT the trainer's graph: blade-volume-train
 * differentiates recorded paths (diff_render.rs:943-1921), which
 * bv_volume_graph_step below restates.  rgba_forward is this batch's
 * bv_trace_rays output, d_rgba the loss gradient [count][4].  Gradients are
 * ADDED to `grads`. */
int bv_adjoint(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
               const float* rgba_forward, const float* d_rgba, const bv_grads* grads, int kernel, void* stream);

/* One training batch of trace_one_ray with an L2 colour loss against
 * target_rgb [count][3]: forward walk, dL/d(rgba) = (2 (rgb - target), 0),
 * the bv_adjoint gradient, ADDED to `grads`; rgba_out [count][4] (may be NULL
 * for BV_KERNEL_FAST) receives the forward colour and loss[0] (may be NULL)
 * accumulates sum (rgb - target)^2.  The L2 training workload's fused kernel:
 * the shape of one trainer step (record + graph forward + backward,
 * diff_render.rs:7181-7225) with trace_one_ray's walk as the differentiated
 * function; the trainer's own graph is bv_volume_graph_step.  BV_KERNEL_FAST runs it as one
 * fused kernel per batch (with a path cache the library allocates per stream
 * on first use); BV_KERNEL_PLAIN as trace, loss and adjoint launches of the
 * plain port (with per-stream library scratch for dL/d(rgba) and, when
 * rgba_out is NULL, the colour). */
int bv_train_step_l2(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                     const float* target_rgb, float* rgba_out, const bv_grads* grads, float* loss, int kernel,
                     void* stream);

/* bv_train_step_l2 with BV_KERNEL_FAST that also writes each ray's visited-cell
 * sequence (path_cells [count][path_len], as bv_trace_outputs.path) from the
 * fused kernel's own forward walk: the same launch, for verification against
 * a CPU walk (no reference counterpart). */
int bv_train_step_l2_traced(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                            const float* target_rgb, float* rgba_out, const bv_grads* grads, float* loss,
                            uint32_t* path_cells, uint32_t path_len, void* stream);

/* bv_train_step_l2 with BV_KERNEL_FAST that allocates nothing, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No reference counterpart.
 * The ray counter and path cache live in a fixed 64 MiB device workspace of
 * the library (one per GPU, held until the process ends); rays whose walks
 * outgrow the cache are replayed by walking them again, and density /
 * position gradients are added straight into `grads` (no packed table), so
 * the gradients equal bv_train_step_l2's up to FP32 summation order and the
 * colours bit for bit.  Every call on a GPU shares that GPU's workspace:
 * calls that overlap on one GPU (other streams, other threads) race and must
 * be serialised by the caller.  At most 2^31 - 1 rays per call. */
int bv_train_step_l2_noalloc(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                             const float* target_rgb, float* rgba_out, const bv_grads* grads, float* loss,
                             void* stream);

/* ---- The trainer's volumetric graph over recorded paths: the forward and
 * backward of blade-volume-train's build_volumetric_graph_with_options
 * (diff_render.rs:943-1921), which meganeura runs on every training step
 * (session.step(), diff_render.rs:7225), for RadFoam and PowerFoam models
 * without oriented surfaces.  The paths come from bv_record_paths /
 * bv_splat_record (or any recorder with the same [P, L] rows). */
#define BV_GRAPH_DT_POSITIONS 0   /* RadFoam: dt from the bisector crossings of (cell, next) (:1327-1473) */
#define BV_GRAPH_DT_RECORDED 1    /* weighted paths with frozen geometry: the recorder's dt (:1553) */
#define BV_GRAPH_DT_LINEARIZED 2  /* weighted paths: dt_ref + recorded-Jacobian tangent - tangent_ref (:1474-1552) */
#define BV_LOSS_L1 0              /* ColorLoss::L1: per channel mean |pixel - label|, summed */
#define BV_LOSS_SMOOTH_L1 1       /* ColorLoss::SmoothL1 (beta 1), mean over pixels and channels */
#define BV_LOSS_L2 2              /* per channel mean squared error, summed (no reference counterpart) */

/* The graph's parameter tables (device).  The reference keeps per-channel
 * SH tables sh_r/g/b [N,1] + sh_*_rest [N,K-1] and exposure_r/g/b [V,1]; here
 * they are the PointCloudModel layouts below (entry [i][3j + c] is channel c
 * of SH component j; exposure [v][c]). */
typedef struct {
  int64_t n_cells;
  int32_t sh_degree;           /* 0..3 */
  int32_t n_views;             /* rows of exposure */
  const float* log_density;    /* [n] raw density parameter (activation below) */
  const float* positions;      /* [n][3] */
  const float* sh;             /* [n][3K] */
  const float* log_radii;      /* [n] (radius = softplus_100(log_radius)); BV_GRAPH_DT_LINEARIZED only */
  const float* exposure;       /* [n_views][3] per-view RGB gain, or NULL (gain 1, no gradient) */
} bv_graph_params;

/* One batch of P pixel rows of L recorded steps (device).  Steps naming a
 * cell outside the model are padding: nothing of the model is read for them
 * and they add nothing. */
typedef struct {
  uint32_t num_pixels;         /* P */
  uint32_t max_steps;          /* L */
  const uint32_t* cells;       /* [P][L] cell_indices */
  const uint32_t* next_cells;  /* [P][L] next_cell_indices */
  const float* dts;            /* [P][L] recorded_dt */
  const float* mask;           /* [P][L] */
  const uint32_t* status;      /* [P] recorder status, or NULL: with it, steps past status & 0x7fffffff are
                                  taken as padding (mask 0), as the recorder's callers leave them */
  const uint32_t* previous_cells;      /* [P][L], BV_GRAPH_DT_LINEARIZED */
  const float* dt_reference_tangent;   /* [P][L], BV_GRAPH_DT_LINEARIZED */
  const float* dt_grad_previous;       /* [P][L][4] d dt / d (xyz, radius) of the previous site */
  const float* dt_grad_current;        /* [P][L][4] ... of the current site */
  const float* dt_grad_next;           /* [P][L][4] ... of the next site */
  const float* ray_origin;     /* [P][3] */
  const float* ray_dir;        /* [P][3] ray_dir_per_pixel (unit) */
  const float* basis;          /* [P][K-1] the inputs basis_1 .. basis_{K-1}, or NULL: sh_basis(ray_dir)
                                  (diff_render.rs:2062-2090) on the device */
  const uint32_t* view_idx;    /* [P] or NULL (view 0); a view >= n_views gets gain 1 and no gradient */
  const float* labels;         /* [P][3] target colours */
  const float* target_alpha;   /* [P], needed when opacity_weight > 0 */
} bv_graph_batch;

typedef struct {
  float softplus_beta;         /* density activation: 0 = ReLU, else softplus with this beta (:227-232) */
  float background[3];         /* background RGB composited behind the pixel */
  uint32_t loss;               /* BV_LOSS_* */
  float opacity_weight;        /* opacity loss weight * mean((opacity - target_alpha)^2), 0 = off */
  uint32_t dt_mode;            /* BV_GRAPH_DT_* */
  uint32_t pad;
} bv_graph_options;

/* Outputs (device, any may be NULL): loss[0] is ADDED to; pixels [P][3] after
 * exposure and background; opacity [P] = sum of weights; dt [P][L] the
 * graph's dt_from_positions (its second output). */
typedef struct {
  float* loss;
  float* pixels;
  float* opacity;
  float* dt;
} bv_graph_outputs;

/* Gradient tables (device), ADDED to; NULL = that table is not trained. */
typedef struct {
  float* d_log_density;        /* [n] */
  float* d_positions;          /* [n][3] (not in BV_GRAPH_DT_RECORDED) */
  float* d_sh;                 /* [n][3K] */
  float* d_log_radii;          /* [n] (BV_GRAPH_DT_LINEARIZED) */
  float* d_exposure;           /* [n_views][3] */
} bv_graph_grads;

/* Replaces the meganeura graph's forward + backward of one training step
 * (build_volumetric_graph_with_options, diff_render.rs:943-1921, evaluated by
 * session.step() at :7225; the Adam update stays the caller's): loss, pixels
 * and the exact reverse-mode gradients of the loss, ReLU derivative 0 at 0
 * and sign(0) = 0.  Asynchronous on `stream`; nothing is allocated. */
int bv_volume_graph_step(const bv_graph_params* params, const bv_graph_batch* batch,
                         const bv_graph_options* options, const bv_graph_outputs* outputs,
                         const bv_graph_grads* grads, void* stream);

/* Relit render of a camera image (the relight viewer's point-light path,
 * relight.wgsl:190-262, composed with the cell walk): rgba_out [H*W][4].
 * Needs a scene created with materials; settings->eval_mode is ignored. */
int bv_relight_camera(bv_scene* scene, const bv_camera* camera, uint32_t width, uint32_t height,
                      uint32_t start_cell, const bv_trace_settings* settings, const bv_lighting* lighting,
                      float* rgba_out, void* stream);
int bv_relight_rays(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                    const bv_lighting* lighting, float* rgba_out, void* stream);

/* ---- PowerFoam splats: every support sphere a ray hits, clipped by the radical
 * planes of its adjacency row (the Cech graph of the supports), composited in
 * (near, cell) order.  Needs radii.  The sphere BVH is built on the device on
 * first use and rebuilt after bv_scene_update.
 *
 * trace_powerfoam_splats (trace.rs:1456-1546) over a batch of rays; outputs as
 * bv_trace_rays (steps = intervals composited, t_end = near + dt of the last,
 * depth_mode = near + dt/2 of the heaviest).  rays->start_cell(s) only seed
 * last_point (BV_NO_CELL counts as 0). */
int bv_splat_trace_rays(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                        const bv_trace_outputs* outputs, void* stream);
/* PowerFoamGpuSplatTracer::dispatch (gpu/powerfoam_splat.rs:249-319) for a
 * camera image: outputs are [H*W] in row-major pixel order. */
int bv_splat_trace_camera(bv_scene* scene, const bv_camera* camera, uint32_t width, uint32_t height,
                          const bv_trace_settings* settings, const bv_trace_outputs* outputs, void* stream);
/* record_powerfoam_splats_jacobians (trace.rs:1414-1443) without Jacobians:
 * rows [count][max_steps] of cell / near / dt for the first max_steps intervals
 * in (near, cell) order (no transmittance early-out); count_out[count] = row
 * length.  Slots past a row's length are left untouched. */
int bv_splat_record(bv_scene* scene, const bv_rays* rays, uint32_t max_steps, float depth, uint32_t* cells,
                    float* nears, float* dts, uint32_t* count_out, void* stream);
/* The PowerFoam path recorder with its Jacobian streams (PathRecorder::dispatch
 * for power cells, record_powerfoam_splats in radfoam_record_paths.wgsl with
 * jacobian_mode 1, :1030-1096; CPU record_powerfoam_splats_jacobians,
 * trace.rs:1414-1443): per ray, the first max_steps intervals in (near, cell)
 * order.  Buffers [count][max_steps] (Jacobians [count][max_steps][4]): */
typedef struct {
  uint32_t* cells;             /* the interval's cell */
  uint32_t* next_cells;        /* the cell across its exit plane (the cell itself at a sphere or far-plane end) */
  uint32_t* previous_cells;    /* the cell across its entry plane (the cell itself at a sphere or near end) */
  float* dts;                  /* interval length, clamped to max_path_dt */
  float* mask;                 /* 1 for every written step, or NULL */
  float* dt_grad_previous;     /* d dt / d (xyz, radius) of the previous cell */
  float* dt_grad_current;      /* ... of the cell (entry and exit spheres or planes, path_interval_jacobian) */
  float* dt_grad_next;         /* ... of the next cell; all three are 0 when dt was clamped */
  float* dt_reference_tangent; /* sum over the three roles of J . (site - origin, radius) */
  uint32_t* status;            /* [count] steps written | BV_PATH_TRUNCATED when more intervals follow */
} bv_splat_path_outputs;
/* Slots past a row's length are left untouched (the reference's caller clears
 * them).  settings->max_path_dt is the trainer's MAX_PATH_DT (50). */
int bv_splat_record_jacobians(bv_scene* scene, const bv_rays* rays, const bv_record_settings* settings,
                              const bv_splat_path_outputs* outputs, void* stream);

/* RecordPathsArgs (gpu/path_record.rs:142-161): one camera slice of a batched
 * path recording, 80 bytes. */
typedef struct {
  bv_camera camera;            /* position, orientation, fov, principal (camera.depth is not read) */
  uint32_t start_point;        /* cell holding camera.position (Voronoi cells; power cells ignore it) */
  uint32_t pixel_offset;       /* first output row; the slice reads pixel_indices at the same rows */
  uint32_t image_pixel_offset; /* added (u32 wrap) to every pixel index of the slice */
  uint32_t image_width;
  uint32_t image_height;
  uint32_t num_pixels;         /* rows of the slice */
  float depth;
  float max_path_dt;
} bv_record_view;

/* PathRecordBuffers (gpu/path_record.rs:614-670) rows [capacity][max_steps]
 * plus optional per-row ray outputs [capacity] for the volumetric graph. */
typedef struct {
  bv_splat_path_outputs rows;  /* Voronoi cells write cells, next_cells, dts, mask (may be NULL) and status;
                                  power cells write every stream (previous_cells and the Jacobian streams
                                  are then required) */
  float* ray_origins;          /* [capacity][3] or NULL: the origin each recorded row walked */
  float* ray_dirs;             /* [capacity][3] or NULL: its unit direction */
  uint32_t* view_index;        /* [capacity] or NULL: index into `views` of the row's slice */
} bv_record_batch_outputs;

/* Replaces PathRecorder::dispatch_batch (gpu/path_record.rs:519-594) with
 * write_pixel_indices (:1194-1218) done by the caller: `pixel_indices`
 * [capacity] is a device array.  Row r of slice v (pixel_offset <= r <
 * pixel_offset + num_pixels) records the ray of pixel p = pixel_indices[r] +
 * image_pixel_offset of v's camera, ray_dir_for_pixel
 * (radfoam_record_paths.wgsl:133-153: ix = p % width, iy = p / width,
 * normalize(qrot(orientation, ((ndc - principal) * tan(fov / 2), 1))) with the
 * recorder's own qrot; tan evaluated once per slice on the host as
 * (float)tan((double)(0.5f * fov))), origin = camera position, from
 * start_point.  Voronoi scenes record as bv_record_paths, power-cell scenes as
 * bv_splat_record_jacobians (max_steps, depth and max_path_dt per slice as in
 * RecordPathsArgs).  Slices must be ordered by disjoint pixel_offset ranges
 * inside [0, capacity) and have a non-zero image size; rows outside every slice
 * are not written.  Consecutive slices with equal depth and max_path_dt share a
 * launch (up to 24 per launch). */
int bv_record_paths_batch(bv_scene* scene, const bv_record_view* views, uint32_t n_views,
                          const uint32_t* pixel_indices, uint32_t capacity, uint32_t max_steps,
                          const bv_record_batch_outputs* outputs, void* stream);
/* Relit splats (the relight viewer's point-light shading, relight.wgsl:190-262,
 * composed with the PowerFoam splat tracer): every composited interval is shaded
 * by every light at its entry point (normal = point - site, view = -direction,
 * the cell's material); with lighting->visibility each light's sum is scaled by
 * the transmittance of a splat opacity trace from the ray's depth point toward
 * the light.  Needs materials; settings->eval_mode is ignored. */
int bv_splat_relight_camera(bv_scene* scene, const bv_camera* camera, uint32_t width, uint32_t height,
                            const bv_trace_settings* settings, const bv_lighting* lighting, float* rgba_out,
                            void* stream);
int bv_splat_relight_rays(bv_scene* scene, const bv_rays* rays, const bv_trace_settings* settings,
                          const bv_lighting* lighting, float* rgba_out, void* stream);

/* ---- Gaussian ray tracing: the 3DGRT blend of trace_gaussians (trace.rs:695-773)
 * that GaussianGpuCloud + gaussian_trace.wgsl render (gpu/gaussian.rs:94-280,
 * shaders/gaussian_trace.wgsl:66-177), over a device BVH of each particle's
 * support ellipsoid instead of hardware ray queries, and its adjoint. */
typedef struct bv_gaussians bv_gaussians;

/* A Gaussian PointCloudModel (lib.rs:173-239 with Transforms) as host arrays;
 * min_opacity is InitParameters::min_opacity, the response at which each
 * particle's support (its BVH box) ends. */
typedef struct {
  int64_t count;
  const float* means;        /* [n][3] */
  const float* opacity;      /* [n] (points[i].w) */
  const float* rotations;    /* [n][4] quaternion x, y, z, w */
  const float* scales;       /* [n][3], > 0 */
  const float* sh;           /* [n][3K], K = (sh_degree+1)^2, RGB interleaved per component */
  int32_t sh_degree;         /* 0..3 */
  float min_opacity;
} bv_gaussians_desc;

/* GaussianTraceSettings (trace.rs:106-126); min_opacity must be at least the
 * scene's build min_opacity. */
typedef struct {
  float min_opacity;
  float min_transmittance;
  float t_start, t_end;
} bv_gaussian_settings;

/* GaussianTraceResult (trace.rs:128-134) as device outputs. */
typedef struct {
  float* rgba;               /* [count][4], required */
  uint32_t* hits;            /* [count] composited responses, or NULL */
  uint32_t* order;           /* [count][order_len] composited particle indices in order, or NULL */
  uint32_t order_len;
  uint32_t pad;
  uint32_t* stats;           /* [count][2] algorithmic bytes and composited responses, or NULL */
} bv_gaussian_outputs;

/* Gradient tables in device memory (ADDED to): the stored parameters. */
typedef struct {
  float* d_mean;             /* [n][3] */
  float* d_scale;            /* [n][3] */
  float* d_rotation;         /* [n][4] */
  float* d_opacity;          /* [n] */
  float* d_sh;               /* [n][3K] */
} bv_gaussian_grads;

/* Replaces GaussianGpuCloud::new (gpu/gaussian.rs:94): uploads the particles and
 * builds the support BVH.  Synchronous. */
int bv_gaussians_create(const bv_gaussians_desc* desc, bv_gaussians** out);
int bv_gaussians_destroy(bv_gaussians* g);
int bv_gaussians_get_info(const bv_gaussians* g, int64_t* count, int64_t* device_bytes);
/* New parameter values from DEVICE buffers (any may be NULL); geometry changes
 * rebuild the BVH. */
int bv_gaussians_update(bv_gaussians* g, const float* means, const float* opacity, const float* rotations,
                        const float* scales, const float* sh, void* stream);
/* trace_gaussians over a batch of rays (directions need not be unit length). */
int bv_gaussians_trace_rays(bv_gaussians* g, const bv_rays* rays, const bv_gaussian_settings* settings,
                            const bv_gaussian_outputs* out, void* stream);
/* One ray per pixel of a camera (the Gaussian viewer's draw, gaussian_trace.wgsl
 * called per fragment); outputs are [H*W] in row-major pixel order. */
int bv_gaussians_trace_camera(bv_gaussians* g, const bv_camera* camera, uint32_t width, uint32_t height,
                              const bv_gaussian_settings* settings, const bv_gaussian_outputs* out, void* stream);
/* Adjoint of bv_gaussians_trace_rays for dL/d(rgba) given the forward rgba:
 * replays each ray's response order and ADDS FP32 gradients of the stored
 * parameters (no reference API: the trainer differentiates this blend through
 * its meganeura graph, gaussian_splat.rs:2107-2160). */
int bv_gaussians_adjoint(bv_gaussians* g, const bv_rays* rays, const bv_gaussian_settings* settings,
                         const float* rgba_fwd, const float* d_rgba, const bv_gaussian_grads* grads, void* stream);

/* ---- Serving loop (no reference counterpart as an API: the viewer renders one
 * frame per loop turn, blade-volume-view/src/bin/view.rs:593-840).  Resident
 * scenes, view requests admitted while their output image fits the result
 * budget, and continuous batching: one persistent kernel takes 8x8 ray tiles
 * of all admitted requests from a ring that the server's scheduler thread
 * refills while it runs, so a request arriving while others run has its tiles
 * among theirs within one scheduler turn.  The kernel writes each image into a
 * result arena of pinned host memory, and the server copies it to the
 * request's buffer.  A request's image is the same alone and in any mix. */
typedef struct bv_server bv_server;

enum { BV_RENDER_WALK = 0, BV_RENDER_SPLAT = 1 };

typedef struct {
  int64_t hbm_budget_bytes;  /* result arena for admitted requests' output images (pinned host memory
                                the kernel writes; admission waits for room in it) */
  uint32_t max_active;       /* admitted requests at once (1..4095) */
  uint32_t tiles_ahead;      /* 8x8 tiles the scheduler keeps published and unfinished
                                (0 = the kernel's resident blocks and a quarter more; at most 32768) */
} bv_server_config;

typedef struct {
  uint32_t scene;            /* id from bv_server_add_scene */
  uint32_t width, height;
  uint32_t mode;             /* BV_EVAL_SH or BV_EVAL_RELIGHT */
  uint32_t start_cell;       /* walk only; BV_NO_CELL: the camera's containing cell, found at submission */
  uint32_t renderer;         /* BV_RENDER_WALK (cell walk from start_cell) or BV_RENDER_SPLAT
                                (PowerFoam splats, the viewer's PowerFoam path; needs radii) */
  uint32_t pad[2];
  bv_camera camera;
  bv_trace_settings settings;
  bv_lighting lighting;      /* BV_EVAL_RELIGHT only */
  float* host_rgba;          /* [H*W][4] host buffer, written before the request completes */
} bv_view_request;

typedef struct {
  uint64_t submitted, completed, rejected;
  double p50_ms, p99_ms, mean_ms, max_ms;  /* submit -> image in host memory */
  double mrays_per_s;        /* completed rays / (last completion - first submission) */
  double busy_ms;            /* time the scheduler waited on device work */
  int64_t peak_bytes;        /* peak arena use */
} bv_server_stats;

int bv_server_create(const bv_server_config* config, bv_server** out_server);
/* The scene must outlive the server and must not be updated (bv_scene_update)
 * while requests on it are queued or running. */
int bv_server_add_scene(bv_server* server, bv_scene* scene, uint32_t* out_id);
/* Queues a request; BV_ERR_CAPACITY if its image can never fit the budget. */
int bv_server_submit(bv_server* server, const bv_view_request* request, uint64_t* out_ticket);
/* *out_done = 1 once the request's image is in its host buffer. */
int bv_server_poll(bv_server* server, uint64_t ticket, int* out_done);
/* Blocks until the image is in its host buffer, returns its latency and
 * releases the ticket (later calls with it return BV_ERR_INVALID). */
int bv_server_wait(bv_server* server, uint64_t ticket, double* out_latency_ms);
int bv_server_get_stats(bv_server* server, bv_server_stats* out_stats);
/* Drains every queued request, then releases the server (not its scenes). */
int bv_server_destroy(bv_server* server);

/* L2 loss gradient helper for the training workload: d_rgba = 2 (rgb - target),
 * alpha gradient 0, and loss[0] += sum (rgb - target)^2 (loss may be NULL). */
int bv_l2_loss_grad(const float* rgba, const float* target_rgb, int64_t count, float* d_rgba, float* loss,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BV_H */
