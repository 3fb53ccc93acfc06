"""ctypes front end of the CPU oracle (``liboracle.so``, built from bv_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only to check or
to time the reference algorithm on the host.  The product (``blade-volume_amd``)
never imports it and has no CPU fallback.

The functions restate ``blade-volume/src/trace.rs`` (see the header of
bv_oracle.cpp for the per-function file:line map).  Arrays are numpy; the
float32 entry points are the parity oracle, the float64 ones are the twin used
for central finite differences of the adjoint.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

EVAL_OPACITY, EVAL_CONSTANT, EVAL_SH = 0, 1, 2
NO_CELL = 0xFFFFFFFF
TRUNCATED_BIT = 0x80000000


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        # BVO_LIBRARY: another build of the oracle (the sanitizer build)
        alt = os.environ.get("BVO_LIBRARY")
        if alt:
            _lib = C.CDLL(alt)
            _declare(_lib)
            return _lib
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class _SceneF32(C.Structure):
    _fields_ = [("n", C.c_int64), ("xyz", C.c_void_p), ("radius", C.c_void_p), ("density", C.c_void_p),
                ("sh", C.c_void_p), ("sh_degree", C.c_int32), ("offsets", C.c_void_p), ("neighbors", C.c_void_p)]


class _SceneF64(_SceneF32):
    pass


class _SettingsF32(C.Structure):
    _fields_ = [("weight_threshold", C.c_float), ("max_steps", C.c_uint32), ("depth", C.c_float),
                ("eval_mode", C.c_uint32), ("const_rgb", C.c_float * 3)]


class _SettingsF64(C.Structure):
    _fields_ = [("weight_threshold", C.c_double), ("max_steps", C.c_uint32), ("depth", C.c_double),
                ("eval_mode", C.c_uint32), ("const_rgb", C.c_double * 3)]


def _declare(lib):
    vp, i64, u32, f32, i32 = C.c_void_p, C.c_int64, C.c_uint32, C.c_float, C.c_int
    lib.bvo_trace_rays_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]
    lib.bvo_trace_rays_f64.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]
    lib.bvo_trace_objects_f32.argtypes = [vp, vp, vp, vp, u32, i64, vp, vp, vp, vp, i32]
    lib.bvo_trace_objects_f64.argtypes = [vp, vp, vp, vp, u32, i64, vp, vp, vp, vp, i32]
    lib.bvo_record_paths_f32.argtypes = [vp, i64, vp, vp, vp, u32, f32, f32, vp, vp, vp, vp, i32]
    lib.bvo_adjoint_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.bvo_adjoint_f64.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.bvo_adjoint_mt_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, i32]
    lib.bvo_max_threads.argtypes = []
    lib.bvo_max_threads.restype = i32
    lib.bvo_camera_rays_f32.argtypes = [vp, vp, vp, vp, u32, u32, vp, vp]
    lib.bvo_recorder_rays_f32.argtypes = [vp, vp, vp, vp, u32, u32, u32, vp, i64, vp, vp]
    lib.bvo_trace_relight_f32.argtypes = [vp, vp, i64, vp, vp, vp, vp, vp, i32, i32, vp, vp, i32]
    lib.bvo_shade_point_light_f32.argtypes = [vp, vp, vp, vp, vp, f32, vp, vp]
    lib.bvo_containing_cell_f32.argtypes = [vp, vp]
    lib.bvo_containing_cell_f32.restype = u32
    lib.bvo_eval_sh_f32.argtypes = [vp, u32, vp, vp]
    lib.bvo_support_interval_f32.argtypes = [C.c_int, vp, vp, vp, f32, f32, f32, vp]
    lib.bvo_support_interval_f32.restype = C.c_int
    lib.bvo_trace_splats_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32]
    lib.bvo_record_splats_f32.argtypes = [vp, i64, vp, vp, f32, u32, vp, vp, vp, vp, i32]
    lib.bvo_trace_splat_relight_f32.argtypes = [vp, vp, i64, vp, vp, vp, vp, i32, i32, vp, vp, i32]
    lib.bvo_trace_gaussians_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, u32, i32]
    lib.bvo_trace_gaussians_f64.argtypes = [vp, i64, vp, vp, vp, vp, i32]
    lib.bvo_adjoint_gaussians_f32.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.bvo_adjoint_gaussians_f64.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp]


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class _PointLight(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("exponent", C.c_float), ("direction", C.c_float * 3),
                ("visibility", C.c_uint32), ("intensity", C.c_float * 3), ("pad1", C.c_float)]


def _lights(lights):
    arr = (_PointLight * max(1, len(lights)))()
    for i, L in enumerate(lights):
        arr[i].position[:] = [float(x) for x in L["position"]]
        arr[i].direction[:] = [float(x) for x in L.get("direction", (0.0, 0.0, 1.0))]
        arr[i].intensity[:] = [float(x) for x in L["intensity"]]
        arr[i].exponent = float(L.get("exponent", 0.0))
        arr[i].visibility = 1 if L.get("visibility") else 0
    return arr


def shade_point_light(light, point, normal, view, albedo, roughness, f0=(0.04, 0.04, 0.04)) -> np.ndarray:
    """shade_point_light (relight.rs:295-324) at one point."""
    lib = _load()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    p, n, v, a, f0a = f(point), f(normal), f(view), f(albedo), f(f0)
    out = np.zeros(3, np.float32)
    L = _lights([light])
    lib.bvo_shade_point_light_f32(C.byref(L[0]), _p(p), _p(n), _p(v), _p(a), float(roughness), _p(f0a), _p(out))
    return out


def trace_relight(scene: "Scene", material, origins, dirs, start, lights, *, visibility=False,
                  specular_f0=(0.04, 0.04, 0.04), weight_threshold=0.001, max_steps=1024, depth=10_000.0,
                  threads=0) -> np.ndarray:
    """Relit walk: trace_one_ray's walk with every integrated segment shaded by
    shade_point_light per light, lights scaled by shadow-walk transmittance."""
    lib = _load()
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.uint32), (n,)))
    m = np.ascontiguousarray(material, np.float32)
    L = _lights(lights)
    f0 = np.ascontiguousarray(specular_f0, np.float32)
    se = _settings(np.float32, weight_threshold, max_steps, depth, EVAL_SH, (1.0, 1.0, 1.0))
    out = np.zeros((n, 4), np.float32)
    lib.bvo_trace_relight_f32(C.byref(st), _p(m), n, _p(o), _p(d), _p(s), C.byref(se), L, len(lights),
                              int(bool(visibility)), _p(f0), _p(out), int(threads))
    del keep
    return out


@dataclass
class Scene:
    """Host scene in the reference's logical layout (PointCloudModel, lib.rs:173-239):
    xyz [N,3], density [N], sh [N, 3K] (interleaved RGB per component), optional
    radius [N] (PowerFoam), CSR offsets [N+1] / neighbors [E] (u32)."""
    xyz: np.ndarray
    density: np.ndarray
    sh: np.ndarray
    sh_degree: int
    offsets: np.ndarray
    neighbors: np.ndarray
    radius: Optional[np.ndarray] = None

    def as_dtype(self, dtype) -> "Scene":
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=dtype)
        return Scene(c(self.xyz), c(self.density), c(self.sh), int(self.sh_degree),
                     np.ascontiguousarray(self.offsets, np.uint32), np.ascontiguousarray(self.neighbors, np.uint32),
                     c(self.radius))

    def _struct(self, dtype):
        s = self.as_dtype(dtype)
        cls = _SceneF32 if dtype == np.float32 else _SceneF64
        st = cls(len(s.density), _p(s.xyz), _p(s.radius), _p(s.density), _p(s.sh), s.sh_degree, _p(s.offsets),
                 _p(s.neighbors))
        return st, s  # keep the converted arrays alive


def _settings(dtype, weight_threshold, max_steps, depth, eval_mode, const_rgb):
    cls = _SettingsF32 if dtype == np.float32 else _SettingsF64
    arr = (C.c_float * 3) if dtype == np.float32 else (C.c_double * 3)
    return cls(weight_threshold, max_steps, depth, eval_mode, arr(*const_rgb))


def trace_rays(scene: Scene, origins, dirs, start, *, weight_threshold=0.001, max_steps=1024, depth=10_000.0,
               eval_mode=EVAL_SH, const_rgb=(1.0, 1.0, 1.0), dtype=np.float32, threads=0) -> dict:
    """trace_one_ray (trace.rs:1680-1820) over a batch of rays."""
    lib = _load()
    st, keep = scene._struct(dtype)
    o = np.ascontiguousarray(origins, dtype).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype).reshape(-1, 3)
    n = o.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.uint32), (n,)))
    out = dict(rgba=np.zeros((n, 4), dtype), steps=np.zeros(n, np.uint32), last_point=np.zeros(n, np.uint32),
               t_end=np.zeros(n, dtype), depth_mode=np.zeros(n, dtype), peak_weight=np.zeros(n, dtype),
               peak_point=np.zeros(n, np.uint32))
    se = _settings(dtype, weight_threshold, max_steps, depth, eval_mode, const_rgb)
    fn = lib.bvo_trace_rays_f32 if dtype == np.float32 else lib.bvo_trace_rays_f64
    fn(C.byref(st), n, _p(o), _p(d), _p(s), C.byref(se), _p(out["rgba"]), _p(out["steps"]), _p(out["last_point"]),
       _p(out["t_end"]), _p(out["depth_mode"]), _p(out["peak_weight"]), _p(out["peak_point"]), int(threads))
    del keep
    return out


def trace_objects(objects, origins, dirs, *, weight_threshold=0.001, max_steps=1024, eval_mode=EVAL_SH,
                  const_rgb=(1.0, 1.0, 1.0), dtype=np.float32, threads=0) -> np.ndarray:
    """Multi-object scene trace (trace_scene, scene_trace_core.wgsl:406-465) over
    a batch of world rays.  objects: sequence of (Scene, world_to_object [12]
    row-major 3x4, sphere (cx, cy, cz, r), start_cell).  Returns rgba [n, 4]."""
    lib = _load()
    n_obj = len(objects)
    packed = [sc._struct(dtype) for sc, _, _, _ in objects]
    arr = (_SceneF32 * max(n_obj, 1))(*[st for st, _ in packed])
    xf = np.ascontiguousarray(np.asarray([np.asarray(m, np.float64).reshape(12) for _, m, _, _ in objects]
                                         if n_obj else np.zeros((1, 12)), dtype).reshape(-1))
    sp = np.ascontiguousarray(np.asarray([np.asarray(b, np.float64).reshape(4) for _, _, b, _ in objects]
                                         if n_obj else np.zeros((1, 4)), dtype).reshape(-1))
    starts = np.ascontiguousarray([int(st) for _, _, _, st in objects] or [0], np.uint32)
    o = np.ascontiguousarray(origins, dtype).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype).reshape(-1, 3)
    n = o.shape[0]
    rgba = np.zeros((n, 4), dtype)
    se = _settings(dtype, weight_threshold, max_steps, 0.0, eval_mode, const_rgb)
    fn = lib.bvo_trace_objects_f32 if dtype == np.float32 else lib.bvo_trace_objects_f64
    fn(C.byref(arr), _p(xf), _p(sp), _p(starts), n_obj, n, _p(o), _p(d), C.byref(se), _p(rgba), int(threads))
    del packed
    return rgba


def trace_paths(scene: Scene, origins, dirs, start, *, path_len: int, weight_threshold=0.001, max_steps=1024,
                depth=10_000.0, eval_mode=EVAL_SH, threads=0) -> dict:
    """trace_one_ray's visited-cell sequence: path [n, path_len] = the current
    cell of each of the first path_len loop iterations (the stopping one
    included), NO_CELL after; steps [n]."""
    lib = _load()
    if not hasattr(lib, "_paths_declared"):
        vp = C.c_void_p
        lib.bvo_trace_paths_f32.argtypes = [vp, C.c_int64, vp, vp, vp, vp, C.c_uint32, vp, vp, C.c_int]
        lib._paths_declared = True
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.uint32), (n,)))
    se = _settings(np.float32, weight_threshold, max_steps, depth, eval_mode, (1.0, 1.0, 1.0))
    path = np.full((n, path_len), NO_CELL, np.uint32)
    steps = np.zeros(n, np.uint32)
    lib.bvo_trace_paths_f32(C.byref(st), n, _p(o), _p(d), _p(s), C.byref(se), path_len, _p(path), _p(steps),
                            int(threads))
    del keep
    return dict(path=path, steps=steps)


def record_paths(scene: Scene, origins, dirs, start, *, max_steps: int, depth=10_000.0, max_path_dt=float("inf"),
                 threads=0) -> dict:
    """GPU recorder semantics (radfoam_record_paths.wgsl:1220-1355): [P,L] rows
    of cell / next cell / dt and status = count | truncated<<31.  Unwritten
    slots keep the fill values (cells = NO_CELL, dt = nan)."""
    lib = _load()
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.uint32), (n,)))
    cells = np.full((n, max_steps), NO_CELL, np.uint32)
    nxt = np.full((n, max_steps), NO_CELL, np.uint32)
    dts = np.full((n, max_steps), np.nan, np.float32)
    status = np.zeros(n, np.uint32)
    lib.bvo_record_paths_f32(C.byref(st), n, _p(o), _p(d), _p(s), max_steps, depth, max_path_dt, _p(cells), _p(nxt),
                             _p(dts), _p(status), int(threads))
    del keep
    return dict(cells=cells, next_cells=nxt, dts=dts, status=status)


def compact_paths(status, cells, max_steps: int, base: int = 0) -> dict:
    """compact_paths (path_compact.wgsl:24-50, PathCompactor::dispatch
    path_compact.rs:63-89): the records appended at `base` for every path's
    live prefix (status & 0x7fffffff steps of its max_steps-wide row): dense
    slot p * max_steps + step, cell, pixel p, active 1.  Paths in path order
    (the reference's per-path atomic allows any block order; this is the one
    the device stream produces).  Returns the new count and the appended
    records."""
    status = np.asarray(status).astype(np.uint32).reshape(-1)
    n = len(status)
    cells = np.asarray(cells).astype(np.uint32).reshape(n, max_steps)
    steps = np.minimum(status & np.uint32(0x7FFFFFFF), np.uint32(max_steps)).astype(np.int64)
    pixel = np.repeat(np.arange(n, dtype=np.int64), steps)
    first = np.cumsum(steps) - steps
    step = np.arange(int(steps.sum()), dtype=np.int64) - np.repeat(first, steps)
    slots = pixel * max_steps + step
    return dict(count=int(base) + int(steps.sum()), dense_slots=slots.astype(np.uint32),
                cells=cells.reshape(-1)[slots], pixel_indices=pixel.astype(np.uint32),
                active=np.ones(len(slots), np.float32))


def adjoint(scene: Scene, origins, dirs, start, d_rgba, *, weight_threshold=0.001, max_steps=1024, depth=10_000.0,
            dtype=np.float32, threads=None) -> dict:
    """Gradient of trace_one_ray (SH mode, plain RadFoam) given dL/d(rgba).
    threads=None is the serial parity oracle; an integer (0 = all cores)
    runs the FP32 timing baseline over OpenMP threads with atomic adds."""
    lib = _load()
    st, keep = scene._struct(dtype)
    o = np.ascontiguousarray(origins, dtype).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype).reshape(-1, 3)
    n = o.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.uint32), (n,)))
    g = np.ascontiguousarray(d_rgba, dtype).reshape(n, 4)
    N = len(scene.density)
    out = dict(density=np.zeros(N, dtype), sh=np.zeros(scene.sh.size, dtype), xyz=np.zeros(N * 3, dtype))
    se = _settings(dtype, weight_threshold, max_steps, depth, EVAL_SH, (1.0, 1.0, 1.0))
    if threads is not None:
        assert dtype == np.float32, "the threaded adjoint is FP32"
        lib.bvo_adjoint_mt_f32(C.byref(st), n, _p(o), _p(d), _p(s), C.byref(se), _p(g), _p(out["density"]),
                               _p(out["sh"]), _p(out["xyz"]), int(threads))
    else:
        fn = lib.bvo_adjoint_f32 if dtype == np.float32 else lib.bvo_adjoint_f64
        fn(C.byref(st), n, _p(o), _p(d), _p(s), C.byref(se), _p(g), _p(out["density"]), _p(out["sh"]),
           _p(out["xyz"]))
    del keep
    out["sh"] = out["sh"].reshape(scene.sh.shape)
    out["xyz"] = out["xyz"].reshape(N, 3)
    return out


def max_threads() -> int:
    """Threads the oracle's OpenMP loops use with threads=0."""
    return int(_load().bvo_max_threads())


def camera_rays(position, quat, tan_half, principal, width: int, height: int):
    """Pixel-centre rays of radfoam.wgsl:45-52 (qrot of common.wgsl:19-23)."""
    lib = _load()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    pos, q, th, pp = f(position), f(quat), f(tan_half), f(principal)
    o = np.zeros((height * width, 3), np.float32)
    d = np.zeros((height * width, 3), np.float32)
    lib.bvo_camera_rays_f32(_p(pos), _p(q), _p(th), _p(pp), width, height, _p(o), _p(d))
    return o, d


def recorder_tan_half(fov) -> np.ndarray:
    """tan(fov / 2) of one recorder slice as the host evaluates it once per
    slice: float32 half angle, double tangent (C library), one rounding."""
    half = np.float32(0.5) * np.asarray(fov, np.float32)
    return np.array([math.tan(float(h)) for h in half], np.float32)


def recorder_rays(view: dict, pixel_indices):
    """Rays of one recorder slice (radfoam_record_paths.wgsl:133-153 with
    record_paths :1227-1230): origins [n, 3] and unnormalised directions
    [n, 3] for the slice's rows.  `view` holds position, orientation (x, y,
    z, w), fov, principal, image_width, image_height and image_pixel_offset;
    pixel_indices are the slice's rows of the index buffer."""
    lib = _load()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    idx = np.ascontiguousarray(pixel_indices, np.uint32).reshape(-1)
    n = len(idx)
    o = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    pos, q, th, pp = f(view["position"]), f(view["orientation"]), recorder_tan_half(view["fov"]), f(view["principal"])
    lib.bvo_recorder_rays_f32(_p(pos), _p(q), _p(th), _p(pp), int(view["image_width"]), int(view["image_height"]),
                              int(view.get("image_pixel_offset", 0)) & 0xFFFFFFFF, _p(idx), n, _p(o), _p(d))
    return o, d


def unit(d) -> np.ndarray:
    """A direction divided by its float32 length ((x*x + y*y) + z*z, IEEE
    sqrt and division): the walk's normalisation (trace.rs:1704-1706)."""
    d = np.asarray(d, np.float32)
    n2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    return d / np.sqrt(n2)[..., None]


def record_batch(scene: Scene, views, pixel_indices, *, capacity: int, max_steps: int, threads=0) -> dict:
    """PathRecorder::dispatch_batch (gpu/path_record.rs:519-594): each slice
    (pixel_offset, num_pixels, start_point, depth, max_path_dt and the camera
    fields of recorder_rays) records its rows of [capacity, L] buffers with
    the walk recorder (record_paths) for Voronoi cells or the splat recorder
    with Jacobians (record_splat_jacobians) for power cells.  Rows outside
    every slice keep the fill values; ray_origins / ray_dirs (unit) /
    view_index are the per-row rays."""
    idx_all = np.asarray(pixel_indices, np.uint32).reshape(-1)
    power = scene.radius is not None
    L = max_steps
    fields = (("cells", NO_CELL, np.uint32, ()), ("next_cells", NO_CELL, np.uint32, ()),
              ("dts", np.nan, np.float32, ()), ("status", 0, np.uint32, None))
    if power:
        fields += (("previous_cells", NO_CELL, np.uint32, ()), ("dt_grad_previous", np.nan, np.float32, (4,)),
                   ("dt_grad_current", np.nan, np.float32, (4,)), ("dt_grad_next", np.nan, np.float32, (4,)),
                   ("dt_reference_tangent", np.nan, np.float32, ()))
    out = {k: np.full((capacity,) if tail is None else (capacity, L) + tail, fill, dt)
           for k, fill, dt, tail in fields}
    out["ray_origins"] = np.full((capacity, 3), np.nan, np.float32)
    out["ray_dirs"] = np.full((capacity, 3), np.nan, np.float32)
    out["view_index"] = np.full(capacity, NO_CELL, np.uint32)
    for k, v in enumerate(views):
        a, n = int(v["pixel_offset"]), int(v["num_pixels"])
        if n == 0:
            continue
        o, d = recorder_rays(v, idx_all[a:a + n])
        if power:
            r = record_splat_jacobians(scene, o, d, max_steps=L, depth=v["depth"], max_path_dt=v["max_path_dt"],
                                       threads=threads)
        else:
            r = record_paths(scene, o, d, int(v["start_point"]), max_steps=L, depth=v["depth"],
                             max_path_dt=v["max_path_dt"], threads=threads)
        for key in r:
            out[key][a:a + n] = r[key]
        out["ray_origins"][a:a + n] = o
        out["ray_dirs"][a:a + n] = unit(d)
        out["view_index"][a:a + n] = k
    return out


def containing_cell(scene: Scene, position) -> int:
    """lib.rs:526-541 (brute force, lowest index on ties)."""
    lib = _load()
    st, keep = scene._struct(np.float32)
    q = np.ascontiguousarray(position, np.float32)
    r = int(lib.bvo_containing_cell_f32(C.byref(st), _p(q)))
    del keep
    return r


def eval_sh(scene: Scene, cell: int, direction) -> np.ndarray:
    """eval_rgb_sh (trace.rs:397-402) of one cell."""
    lib = _load()
    st, keep = scene._struct(np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    out = np.zeros(3, np.float32)
    lib.bvo_eval_sh_f32(C.byref(st), cell, _p(d), _p(out))
    del keep
    return out


def support_interval(bounded: bool, origin, direction, center, radius, t0, t1):
    """support_interval (trace.rs:209-238); None when the clipped interval is empty."""
    lib = _load()
    f = lambda a: np.ascontiguousarray(a, np.float32)
    ro, rd, c = f(origin), f(direction), f(center)
    out = np.zeros(2, np.float32)
    ok = lib.bvo_support_interval_f32(int(bounded), _p(ro), _p(rd), _p(c), radius, t0, t1, _p(out))
    return (float(out[0]), float(out[1])) if ok else None


def trace_splats(scene: Scene, origins, dirs, start=0, *, weight_threshold=0.001, max_steps=1024, depth=10_000.0,
                 eval_mode=EVAL_SH, const_rgb=(1.0, 1.0, 1.0), threads=0) -> dict:
    """trace_powerfoam_splats (trace.rs:1456-1546): every support sphere the ray
    hits, clipped by its adjacency row's radical planes, composited in
    (near, cell) order.  Exhaustive over cells: O(N) per ray."""
    assert scene.radius is not None, "PowerFoam splats require radii"
    lib = _load()
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(np.asarray(start, np.uint32), (n,)))
    out = dict(rgba=np.zeros((n, 4), np.float32), steps=np.zeros(n, np.uint32), last_point=np.zeros(n, np.uint32),
               t_end=np.zeros(n, np.float32), depth_mode=np.zeros(n, np.float32),
               peak_weight=np.zeros(n, np.float32), peak_point=np.zeros(n, np.uint32))
    se = _settings(np.float32, weight_threshold, max_steps, depth, eval_mode, const_rgb)
    lib.bvo_trace_splats_f32(C.byref(st), n, _p(o), _p(d), _p(s), C.byref(se), _p(out["rgba"]), _p(out["steps"]),
                             _p(out["last_point"]), _p(out["t_end"]), _p(out["depth_mode"]),
                             _p(out["peak_weight"]), _p(out["peak_point"]), int(threads))
    del keep
    return out


def record_splats(scene: Scene, origins, dirs, *, max_steps: int, depth=10_000.0, threads=0) -> dict:
    """record_powerfoam_splats_jacobians (trace.rs:1414-1443) without Jacobians:
    [P, L] rows of (cell, near, dt) in (near, cell) order; count[P].  Unwritten
    slots keep the fill values (cells = NO_CELL, near / dt = nan)."""
    assert scene.radius is not None, "PowerFoam splats require radii"
    lib = _load()
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    cells = np.full((n, max_steps), NO_CELL, np.uint32)
    nears = np.full((n, max_steps), np.nan, np.float32)
    dts = np.full((n, max_steps), np.nan, np.float32)
    count = np.zeros(n, np.uint32)
    lib.bvo_record_splats_f32(C.byref(st), n, _p(o), _p(d), depth, max_steps, _p(cells), _p(nears), _p(dts),
                              _p(count), int(threads))
    del keep
    return dict(cells=cells, nears=nears, dts=dts, count=count)


def _splat_jac_lib():
    lib = _load()
    if not hasattr(lib, "_splat_jac_declared"):
        vp = C.c_void_p
        lib.bvo_record_splat_jac_f32.argtypes = [vp, C.c_int64, vp, vp, C.c_float, C.c_uint32, C.c_float] + [vp] * 9 \
            + [C.c_int]
        lib.bvo_splat_interval_jac_f64.argtypes = [vp, vp, vp, C.c_double, C.c_uint32, C.c_double, vp, vp]
        lib.bvo_splat_interval_jac_f64.restype = C.c_int
        lib._splat_jac_declared = True
    return lib


def record_splat_jacobians(scene: Scene, origins, dirs, *, max_steps: int, depth=10_000.0, max_path_dt=50.0,
                           threads=0) -> dict:
    """The PowerFoam recorder with Jacobian streams (radfoam_record_paths.wgsl,
    jacobian_mode 1; record_powerfoam_splats_jacobians trace.rs:1414-1443):
    [P, L] rows of cell, next, previous, dt (clamped to max_path_dt, zero
    Jacobians when clamped), dt_grad_previous / current / next [P, L, 4] and
    the reference tangent; status = count | bit 31 when more intervals
    follow.  Unwritten slots keep the fill values (NO_CELL / nan)."""
    assert scene.radius is not None, "PowerFoam splats require radii"
    lib = _splat_jac_lib()
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    L = max_steps
    out = dict(cells=np.full((n, L), NO_CELL, np.uint32), next_cells=np.full((n, L), NO_CELL, np.uint32),
               previous_cells=np.full((n, L), NO_CELL, np.uint32), dts=np.full((n, L), np.nan, np.float32),
               dt_grad_previous=np.full((n, L, 4), np.nan, np.float32),
               dt_grad_current=np.full((n, L, 4), np.nan, np.float32),
               dt_grad_next=np.full((n, L, 4), np.nan, np.float32),
               dt_reference_tangent=np.full((n, L), np.nan, np.float32), status=np.zeros(n, np.uint32))
    lib.bvo_record_splat_jac_f32(C.byref(st), n, _p(o), _p(d), depth, L, max_path_dt, _p(out["cells"]),
                                 _p(out["next_cells"]), _p(out["previous_cells"]), _p(out["dts"]),
                                 _p(out["dt_grad_previous"]), _p(out["dt_grad_current"]), _p(out["dt_grad_next"]),
                                 _p(out["dt_reference_tangent"]), _p(out["status"]), int(threads))
    del keep
    return out


def splat_interval_jacobian64(scene: Scene, origin, direction, cell: int, *, depth=10_000.0, max_path_dt=1e30):
    """One cell's splat interval in FP64: None when the ray misses it, else
    dict(near, dt, previous, next, dt_grad_previous / current / next [4])."""
    lib = _splat_jac_lib()
    st, keep = scene._struct(np.float64)
    o = np.ascontiguousarray(origin, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    out = np.zeros(14, np.float64)
    ids = np.zeros(2, np.uint32)
    ok = lib.bvo_splat_interval_jac_f64(C.byref(st), _p(o), _p(d), depth, int(cell), max_path_dt, _p(out), _p(ids))
    del keep
    if not ok:
        return None
    return dict(near=out[0], dt=out[1], previous=int(ids[0]), next=int(ids[1]), dt_grad_previous=out[2:6],
                dt_grad_current=out[6:10], dt_grad_next=out[10:14])


def trace_splat_relight(scene: Scene, material, origins, dirs, lights, *, visibility=False,
                        specular_f0=(0.04, 0.04, 0.04), weight_threshold=0.001, max_steps=1024, depth=10_000.0,
                        threads=0) -> np.ndarray:
    """Relit splats: trace_powerfoam_splats' intervals shaded by shade_point_light
    per light at each interval's entry point, lights scaled by the transmittance
    of a splat opacity trace from the depth point toward the light."""
    lib = _load()
    st, keep = scene._struct(np.float32)
    o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = o.shape[0]
    m = np.ascontiguousarray(material, np.float32)
    L = _lights(lights)
    f0 = np.ascontiguousarray(specular_f0, np.float32)
    se = _settings(np.float32, weight_threshold, max_steps, depth, EVAL_SH, (1.0, 1.0, 1.0))
    out = np.zeros((n, 4), np.float32)
    lib.bvo_trace_splat_relight_f32(C.byref(st), _p(m), n, _p(o), _p(d), C.byref(se), L, len(lights),
                                    int(bool(visibility)), _p(f0), _p(out), int(threads))
    del keep
    return out


class _Gaussians(C.Structure):
    _fields_ = [("n", C.c_int64), ("mean", C.c_void_p), ("opacity", C.c_void_p), ("rot", C.c_void_p),
                ("scale", C.c_void_p), ("sh", C.c_void_p), ("sh_degree", C.c_int32)]


@dataclass
class GaussianScene:
    """A Gaussian PointCloudModel (lib.rs:173-239 with Transforms): mean [N,3],
    opacity [N] (points[i].w), rotation [N,4] quaternion (x, y, z, w), scale
    [N,3], SH [N, 3K] interleaved RGB per component."""
    mean: np.ndarray
    opacity: np.ndarray
    rotation: np.ndarray
    scale: np.ndarray
    sh: np.ndarray
    sh_degree: int

    def _struct(self, dtype):
        c = lambda a: np.ascontiguousarray(a, dtype)
        keep = [c(self.mean), c(self.opacity), c(self.rotation), c(self.scale), c(self.sh)]
        st = _Gaussians(len(keep[1]), *[_p(a) for a in keep], int(self.sh_degree))
        return st, keep


def trace_gaussians(scene: GaussianScene, origins, dirs, *, min_opacity=0.01, min_transmittance=0.01,
                    t_start=0.0, t_end=10_000.0, record: int = 0, dtype=np.float32, threads=0) -> dict:
    """trace_gaussians (trace.rs:695-773) over a batch of rays.  With record > 0
    the first `record` composited particle indices per ray are returned in
    order (rows padded with NO_CELL)."""
    lib = _load()
    st, keep = scene._struct(dtype)
    o = np.ascontiguousarray(origins, dtype).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype).reshape(-1, 3)
    n = o.shape[0]
    s4 = np.array([min_opacity, min_transmittance, t_start, t_end], dtype)
    rgba = np.zeros((n, 4), dtype)
    if dtype == np.float32:
        hits = np.zeros(n, np.uint32)
        order = np.full((n, max(record, 1)), NO_CELL, np.uint32)
        lib.bvo_trace_gaussians_f32(C.byref(st), n, _p(o), _p(d), _p(s4), _p(rgba), _p(hits),
                                    _p(order) if record else None, int(record), int(threads))
        del keep
        out = dict(rgba=rgba, hits=hits)
        if record:
            out["order"] = order
        return out
    lib.bvo_trace_gaussians_f64(C.byref(st), n, _p(o), _p(d), _p(s4), _p(rgba), int(threads))
    del keep
    return dict(rgba=rgba)


def adjoint_gaussians(scene: GaussianScene, origins, dirs, d_rgba, *, min_opacity=0.01, min_transmittance=0.01,
                      t_start=0.0, t_end=10_000.0, dtype=np.float32) -> dict:
    """Gradient of trace_gaussians' rgba w.r.t. mean, scale, rotation (the
    stored quaternion components), opacity and SH, given dL/d(rgba); the hit
    set and order are held fixed."""
    lib = _load()
    st, keep = scene._struct(dtype)
    o = np.ascontiguousarray(origins, dtype).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dtype).reshape(-1, 3)
    n = o.shape[0]
    g = np.ascontiguousarray(d_rgba, dtype).reshape(n, 4)
    N = len(scene.opacity)
    out = dict(mean=np.zeros((N, 3), dtype), scale=np.zeros((N, 3), dtype), rotation=np.zeros((N, 4), dtype),
               opacity=np.zeros(N, dtype), sh=np.zeros(scene.sh.shape, dtype))
    s4 = np.array([min_opacity, min_transmittance, t_start, t_end], dtype)
    fn = lib.bvo_adjoint_gaussians_f32 if dtype == np.float32 else lib.bvo_adjoint_gaussians_f64
    fn(C.byref(st), n, _p(o), _p(d), _p(s4), _p(g), _p(out["mean"]), _p(out["scale"]), _p(out["rotation"]),
       _p(out["opacity"]), _p(out["sh"]))
    del keep
    return out


def of_host(h):
    """The oracle's scene type holding the arrays of a host scene
    (bv.scenes.HostScene, or GaussianHostScene: told apart by its `mean`)."""
    if hasattr(h, "mean"):
        return GaussianScene(h.mean, h.opacity, h.rotation, h.scale, h.sh, h.sh_degree)
    return Scene(h.xyz, h.density, h.sh, h.sh_degree, h.offsets, h.neighbors, h.radius)


# ---------------------------------------------------------------- trainer graph
# The volumetric graph of blade-volume-train (diff_render.rs:943-1921), forward
# and reverse mode per pixel row (bv_graph_oracle.cpp).

DT_POSITIONS, DT_RECORDED, DT_LINEARIZED = 0, 1, 2
LOSS_L1, LOSS_SMOOTH_L1, LOSS_L2 = 0, 1, 2


def _graph_structs(T):
    ct = C.c_float if T == np.float32 else C.c_double
    vp = C.c_void_p

    class Params(C.Structure):
        _fields_ = [("n", C.c_int64), ("sh_degree", C.c_int32), ("n_views", C.c_int32), ("log_density", vp),
                    ("positions", vp), ("sh", vp), ("log_radii", vp), ("exposure", vp)]

    class Batch(C.Structure):
        _fields_ = [("P", C.c_int64), ("L", C.c_int32), ("pad", C.c_int32), ("cells", vp), ("next", vp),
                    ("dt_rec", vp), ("mask", vp), ("status", vp), ("prev", vp), ("ref_tangent", vp),
                    ("jac_prev", vp), ("jac_cur", vp), ("jac_next", vp), ("origin", vp), ("dir", vp),
                    ("basis", vp), ("view", vp), ("labels", vp), ("target_alpha", vp)]

    class Options(C.Structure):
        _fields_ = [("softplus_beta", ct), ("background", ct * 3), ("loss", C.c_int32), ("opacity_weight", ct),
                    ("dt_mode", C.c_int32)]

    class Grads(C.Structure):
        _fields_ = [("d_log_density", vp), ("d_positions", vp), ("d_sh", vp), ("d_log_radii", vp),
                    ("d_exposure", vp)]

    return Params, Batch, Options, Grads


_GRAPH_STRUCTS: dict = {}


def graph_step(params: dict, batch: dict, *, softplus_beta=0.0, background=(0.0, 0.0, 0.0), loss=LOSS_L1,
               opacity_weight=0.0, dt_mode=DT_POSITIONS, dtype=np.float32) -> dict:
    """Loss, pixels, opacity, dt_from_positions and the gradients of the trainer
    graph for one batch.  params: log_density [N], positions [N,3], sh [N,3K],
    sh_degree, optional log_radii [N], exposure [V,3].  batch: cells, next_cells,
    dts, mask [P,L]; optional status [P]; previous_cells, dt_reference_tangent
    [P,L] and dt_grad_previous / _current / _next [P,L,4] (weighted paths);
    ray_origin, ray_dir [P,3]; optional basis [P,K-1], view_idx [P];
    labels [P,3]; optional target_alpha [P]."""
    lib = _load()
    if not hasattr(lib, "_graph_declared"):
        for S, T in (("f32", np.float32), ("f64", np.float64)):
            Pm, Bt, Op, Gr = _GRAPH_STRUCTS.setdefault(T, _graph_structs(T))
            fn = getattr(lib, f"bvo_graph_step_{S}")
            fn.argtypes = [C.POINTER(Pm), C.POINTER(Bt), C.POINTER(Op), C.POINTER(Gr), C.c_void_p, C.c_void_p,
                           C.c_void_p]
            fn.restype = C.c_float if T == np.float32 else C.c_double
        lib._graph_declared = True
    T = np.dtype(dtype).type
    Pm, Bt, Op, Gr = _GRAPH_STRUCTS[T]
    keep = []

    def arr(a, dt=T):
        if a is None:
            return None
        x = np.ascontiguousarray(a, dt)
        keep.append(x)
        return x.ctypes.data

    deg = int(params["sh_degree"])
    K = (deg + 1) ** 2
    n = int(np.asarray(params["log_density"]).shape[0])
    cells = np.asarray(batch["cells"])
    P, L = cells.shape
    views = params.get("exposure")
    n_views = 0 if views is None else int(np.asarray(views).shape[0])
    pm = Pm(n, deg, n_views, arr(params["log_density"]), arr(params["positions"]), arr(params["sh"]),
            arr(params.get("log_radii")), arr(views))
    u32 = lambda k: arr(batch.get(k), np.uint32)
    bt = Bt(P, L, 0, u32("cells"), u32("next_cells"), arr(batch["dts"]), arr(batch["mask"]), u32("status"),
            u32("previous_cells"), arr(batch.get("dt_reference_tangent")), arr(batch.get("dt_grad_previous")),
            arr(batch.get("dt_grad_current")), arr(batch.get("dt_grad_next")), arr(batch["ray_origin"]),
            arr(batch["ray_dir"]), arr(batch.get("basis")), u32("view_idx"), arr(batch["labels"]),
            arr(batch.get("target_alpha")))
    op = Op()
    op.softplus_beta = T(softplus_beta)
    for c in range(3):
        op.background[c] = float(background[c])
    op.loss = int(loss)
    op.opacity_weight = T(opacity_weight)
    op.dt_mode = int(dt_mode)
    out = dict(d_log_density=np.zeros(n, T), d_positions=np.zeros((n, 3), T), d_sh=np.zeros((n, 3 * K), T),
               d_log_radii=np.zeros(n, T), d_exposure=np.zeros((max(n_views, 1), 3), T))
    gr = Gr(out["d_log_density"].ctypes.data, out["d_positions"].ctypes.data, out["d_sh"].ctypes.data,
            out["d_log_radii"].ctypes.data if params.get("log_radii") is not None else None,
            out["d_exposure"].ctypes.data if views is not None else None)
    pixels = np.zeros((P, 3), T)
    opacity = np.zeros(P, T)
    dt = np.zeros((P, L), T)
    fn = getattr(lib, "bvo_graph_step_f32" if T == np.float32 else "bvo_graph_step_f64")
    out["loss"] = float(fn(C.byref(pm), C.byref(bt), C.byref(op), C.byref(gr), pixels.ctypes.data,
                           opacity.ctypes.data, dt.ctypes.data))
    del keep
    out.update(pixels=pixels, opacity=opacity, dt=dt)
    return out
