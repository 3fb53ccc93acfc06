"""blade-volume's ray-marching hot path on MI355X: host-side mirror of the
reference's renderer / recorder / trainer interface over the C ABI (include/bv.h).

Reference interface -> here:
  RadFoamGpuCloud::new / containing_point / deinit  (gpu/radfoam.rs:119-137, :676-678, :535-543)
        -> RadFoamGpuCloud
  RadFoamTraceSettings (gpu/radfoam_trace.rs:4-23)   -> RadFoamTraceSettings
  RadFoamGpuTracer::new / dispatch (gpu/radfoam_trace.rs:63-135)
        -> RadFoamGpuTracer (output: [H, W, 4] FP32 device tensor instead of an rgba16float texture)
  trace::trace_one_ray over a batch (trace.rs:1680-1820) -> trace_rays
  PathRecorder::dispatch + PathRecordBuffers (gpu/path_record.rs:433-495, :614-670) -> PathRecorder
  meganeura graph backward (diff_render.rs:943-1921)   -> adjoint / RadFoamRender (torch.autograd.Function)
  GaussianGpuCloud::new (gpu/gaussian.rs:94-280), trace_gaussians (trace.rs:695-773)
        -> GaussianGpuCloud, gaussian_trace_rays, GaussianTracer (+ gaussian_adjoint, GaussianRender)
  PowerFoamGpuSplatTracer (gpu/powerfoam_splat.rs:36-330), trace_powerfoam_splats /
  record_powerfoam_splats_jacobians (trace.rs:1414-1546)
        -> PowerFoamGpuSplatTracer, splat_trace_rays, splat_record (+ splat_relight_*)
  RadFoamGpuDepthTracer (gpu/radfoam_depth.rs:39-124), PowerFoamGpuDepthTracer (gpu/powerfoam_depth.rs:39-210)
        -> RadFoamGpuDepthTracer, PowerFoamGpuDepthTracer ([H, W, 4] FP32 (depth_mode, alpha, peak_weight, 1))
  PathCompactor / PathCompactBuffers (gpu/path_compact.rs:39-230) -> PathCompactor, PathCompactBuffers

Device memory and streams come from torch (plumbing only); every kernel is the
HIP library's.  There is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Dict, Iterable, Optional

import numpy as np

from . import _lib
from ._lib import (EVAL_CONSTANT, EVAL_OPACITY, EVAL_RELIGHT, EVAL_SH, KERNEL_FAST, KERNEL_PLAIN, NO_CELL,
                   PATH_TRUNCATED, RENDER_SPLAT, RENDER_WALK, BvError, check)
from .scenes import Camera, HostScene

__all__ = [
    "RadFoamGpuCloud", "RadFoamGpuTracer", "RadFoamTraceSettings", "TraceSettings", "PathRecorder", "trace_rays",
    "adjoint", "train_step_l2", "l2_loss_grad", "zero_grads", "RadFoamRender", "relight_camera", "relight_rays",
    "lighting", "Server", "EVAL_RELIGHT", "camera_rays", "BvError", "KERNEL_FAST", "KERNEL_PLAIN",
    "EVAL_OPACITY", "EVAL_CONSTANT", "EVAL_SH", "NO_CELL", "PATH_TRUNCATED", "PowerFoamGpuSplatTracer",
    "splat_trace_rays", "splat_record", "splat_relight_camera", "splat_relight_rays", "RENDER_WALK", "RENDER_SPLAT",
    "GaussianGpuCloud", "GaussianTraceSettings", "gaussian_trace_rays", "gaussian_adjoint", "GaussianTracer",
    "GaussianRender", "RadFoamDepthSettings", "RadFoamGpuDepthTracer", "PowerFoamGpuDepthTracer",
    "PathCompactor", "PathCompactBuffers",
]


def _torch():
    import torch

    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _host(a, dtype):
    return np.ascontiguousarray(a, dtype)


def compute_cech(points, radii, stream=None) -> tuple:
    """adjacency::compute_cech (adjacency.rs:406-425) with the default config,
    built on the device: the Cech graph of weighted sites as a CSR
    (offsets [n + 1], neighbors [offsets[n]], both uint32 host arrays, rows
    ascending), an edge {i, j} whenever the FP32 squared distance is
    <= (r_i + r_j)^2 with radii clamped below at 0.  `points` is [n, 3] (or
    [n, 4], xyz taken)."""
    lib = _lib.load()
    xyz = _host(np.asarray(points, np.float32)[:, :3], np.float32)
    r = _host(radii, np.float32).reshape(-1)
    if len(r) != len(xyz):
        raise ValueError(f"compute_cech: radii.len() ({len(r)}) must equal points.len() ({len(xyz)})")
    handle = C.c_void_p()
    check(lib.bv_compute_cech(xyz.ctypes.data if len(xyz) else None, r.ctypes.data if len(r) else None, len(xyz),
                              _stream_ptr(stream), C.byref(handle)))
    try:
        n = C.c_int64()
        e = C.c_uint64()
        check(lib.bv_adjacency_get_info(handle, C.byref(n), C.byref(e)))
        offsets = np.empty(n.value + 1, np.uint32)
        neighbors = np.empty(e.value, np.uint32)
        check(lib.bv_adjacency_read(handle, offsets.ctypes.data, neighbors.ctypes.data if e.value else None))
    finally:
        check(lib.bv_adjacency_destroy(handle))
    return offsets, neighbors


def compute_delaunay(points) -> tuple:
    """adjacency::try_compute_adjacency (adjacency.rs:142-197) with the default
    config: the Delaunay edge graph of the sites as a CSR (offsets [n + 1],
    neighbors [offsets[n]], uint32 host arrays, rows ascending and self-free),
    built by the library's exact incremental tetrahedralization on the host.
    `points` is [n, 3] (or [n, 4], xyz taken)."""
    lib = _lib.load()
    xyz = _host(np.asarray(points, np.float32)[:, :3], np.float32)
    handle = C.c_void_p()
    check(lib.bv_compute_delaunay(xyz.ctypes.data, len(xyz), C.byref(handle)))
    try:
        n = C.c_int64()
        e = C.c_uint64()
        check(lib.bv_adjacency_get_info(handle, C.byref(n), C.byref(e)))
        offsets = np.empty(n.value + 1, np.uint32)
        neighbors = np.empty(e.value, np.uint32)
        check(lib.bv_adjacency_read(handle, offsets.ctypes.data, neighbors.ctypes.data if e.value else None))
    finally:
        check(lib.bv_adjacency_destroy(handle))
    return offsets, neighbors


class RadFoamGpuCloud:
    """Device copy of a PointCloudModel (gpu/radfoam.rs:25-49).  Holds the
    reference layout (walked by the plain port) and the per-edge plane layout
    (walked by the fast kernels)."""

    def __init__(self, scene: HostScene):
        self._lib = _lib.load()
        scene.validate()
        self.host = scene
        keep = [_host(scene.xyz, np.float32), _host(scene.density, np.float32), _host(scene.sh, np.float32),
                _host(scene.offsets, np.uint32), _host(scene.neighbors, np.uint32)]
        rad = None if scene.radius is None else _host(scene.radius, np.float32)
        mat = None if scene.material is None else _host(scene.material, np.float32)
        desc = _lib.SceneDesc(scene.n, keep[0].ctypes.data, None if rad is None else rad.ctypes.data,
                              keep[1].ctypes.data, keep[2].ctypes.data, scene.sh_degree, keep[3].ctypes.data,
                              keep[4].ctypes.data, 0, None if mat is None else mat.ctypes.data)
        handle = C.c_void_p()
        check(self._lib.bv_scene_create(C.byref(desc), C.byref(handle)))
        self.handle = handle
        info = _lib.SceneInfo()
        check(self._lib.bv_scene_get_info(handle, C.byref(info)))
        self.n_sites = info.n_sites
        self.n_edges = info.n_edges
        self.sh_degree = info.sh_degree
        self.bounded = bool(info.bounded)
        self.max_degree = info.max_degree
        self.device_bytes = info.device_bytes
        self.K = (self.sh_degree + 1) ** 2

    def containing_point(self, position) -> int:
        """gpu/radfoam.rs:676-678: exact (power-)nearest site, lowest index on ties."""
        q = _host(position, np.float32)
        out = C.c_uint32()
        check(self._lib.bv_containing_cell(self.handle, q.ctypes.data, C.byref(out), _stream_ptr()))
        return int(out.value)

    def update(self, xyz=None, density=None, sh=None, stream=None) -> None:
        """Copy new parameter values from device tensors and rebuild the layouts."""
        for t in (xyz, density, sh):
            if t is not None:
                assert t.is_cuda and t.dtype == _torch().float32 and t.is_contiguous()
        check(self._lib.bv_scene_update(self.handle, _ptr(xyz), _ptr(density), _ptr(sh), _stream_ptr(stream)))

    def deinit(self) -> None:
        if getattr(self, "handle", None):
            check(self._lib.bv_scene_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.deinit()
        except Exception:  # interpreter shutdown
            pass


@dataclass
class TraceSettings:
    """trace.rs:50-69 (start_point is per ray batch here)."""
    weight_threshold: float = 0.001
    max_steps: int = 1024
    depth: float = 10_000.0
    eval_mode: int = EVAL_SH
    const_rgb: tuple = (1.0, 1.0, 1.0)

    def c(self) -> _lib.TraceSettings:
        return _lib.TraceSettings(self.weight_threshold, self.max_steps, self.depth, self.eval_mode,
                                  (C.c_float * 3)(*self.const_rgb), 0)


@dataclass
class RadFoamTraceSettings:
    """gpu/radfoam_trace.rs:4-23 (defaults max_steps 1024, weight_threshold 0.001)."""
    max_steps: int = 1024
    weight_threshold: float = 0.001


def _rays(origins, dirs, start_cells, start_cell: int = 0) -> _lib.Rays:
    torch = _torch()
    n = origins.shape[0]
    assert origins.is_cuda and dirs.is_cuda and origins.dtype == torch.float32 and dirs.dtype == torch.float32
    assert origins.is_contiguous() and dirs.is_contiguous() and tuple(origins.shape) == (n, 3) == tuple(dirs.shape)
    if start_cells is not None:
        assert start_cells.is_cuda and start_cells.dtype == torch.int32 and start_cells.shape == (n,)
    return _lib.Rays(origins.data_ptr(), dirs.data_ptr(), _ptr(start_cells), start_cell, 0, n)


_OUT_FIELDS = ("rgba", "steps", "last_point", "t_end", "depth_mode", "peak_weight", "peak_point", "stats")


def _alloc_outputs(n: int, fields: Iterable[str], device) -> Dict[str, "object"]:
    torch = _torch()
    out = {}
    for f in fields:
        if f == "rgba":
            out[f] = torch.empty((n, 4), dtype=torch.float32, device=device)
        elif f == "stats":
            out[f] = torch.empty((n, 2), dtype=torch.int32, device=device)
        elif f in ("steps", "last_point", "peak_point"):
            out[f] = torch.empty(n, dtype=torch.int32, device=device)
        elif f in _OUT_FIELDS:
            out[f] = torch.empty(n, dtype=torch.float32, device=device)
        else:
            raise ValueError(f"unknown output {f}")
    return out


def _outputs_c(out: Dict) -> _lib.TraceOutputs:
    path = out.get("path")
    return _lib.TraceOutputs(*[_ptr(out.get(f)) for f in _OUT_FIELDS], _ptr(path),
                             0 if path is None else int(path.shape[1]), 0)


def trace_rays(cloud: RadFoamGpuCloud, origins, dirs, start_cells=None, start_cell: int = 0,
               settings: TraceSettings = TraceSettings(), kernel: int = KERNEL_FAST, outputs=("rgba",),
               out: Optional[Dict] = None, path_len: int = 0, stream=None) -> Dict:
    """trace_one_ray over a batch of device rays; start_cells is an int32 device
    tensor (u32 bits) or None to start every ray in start_cell.  path_len > 0
    adds out["path"] [n, path_len] (int32 holding u32 bits, -1 where unwritten):
    the visited-cell sequence of each walk."""
    n = origins.shape[0]
    if out is None:
        out = _alloc_outputs(n, outputs, origins.device)
        if path_len:
            out["path"] = _torch().full((n, path_len), -1, dtype=_torch().int32, device=origins.device)
    rays = _rays(origins, dirs, start_cells, start_cell)
    check(cloud._lib.bv_trace_rays(cloud.handle, C.byref(rays), C.byref(settings.c()), C.byref(_outputs_c(out)),
                                   kernel, _stream_ptr(stream)))
    return out


def _camera_c(cam: Camera) -> _lib.CameraParams:
    c = _lib.CameraParams()
    c.position[:] = [float(x) for x in cam.position]
    c.depth = float(cam.depth)
    c.orientation[:] = [float(x) for x in cam.orientation]
    c.fov[:] = [float(x) for x in cam.fov]
    c.principal[:] = [float(x) for x in cam.principal]
    return c


def camera_rays(cam: Camera, width: int, height: int, device="cuda"):
    """The rays RadFoamGpuTracer walks (pixel centres), as [H*W, 3] device tensors."""
    torch = _torch()
    lib = _lib.load()
    o = torch.empty((height * width, 3), dtype=torch.float32, device=device)
    d = torch.empty((height * width, 3), dtype=torch.float32, device=device)
    check(lib.bv_camera_rays(C.byref(_camera_c(cam)), width, height, o.data_ptr(), d.data_ptr(), _stream_ptr()))
    return o, d


class RadFoamGpuTracer:
    """gpu/radfoam_trace.rs:38-171: traces one ray per pixel of a camera."""

    def __init__(self, cloud: RadFoamGpuCloud, settings: RadFoamTraceSettings = RadFoamTraceSettings(),
                 kernel: int = KERNEL_FAST):
        self.cloud = cloud
        self.settings = settings
        self.kernel = kernel

    def set_max_steps(self, max_steps: int) -> None:
        self.settings.max_steps = max_steps

    def set_weight_threshold(self, weight_threshold: float) -> None:
        self.settings.weight_threshold = weight_threshold

    def dispatch(self, camera: Camera, size, out=None, start_cell: Optional[int] = None, outputs=("rgba",),
                 stream=None) -> Dict:
        """Render a (width, height) image; out['rgba'] is [H*W, 4] in row-major pixel order."""
        w, h = int(size[0]), int(size[1])
        if out is None:
            out = _alloc_outputs(w * h, outputs, "cuda")
        st = TraceSettings(self.settings.weight_threshold, self.settings.max_steps, float(camera.depth), EVAL_SH)
        sc = NO_CELL if start_cell is None else int(start_cell)
        check(self.cloud._lib.bv_trace_camera(self.cloud.handle, C.byref(_camera_c(camera)), w, h, sc,
                                              C.byref(st.c()), C.byref(_outputs_c(out)), self.kernel,
                                              _stream_ptr(stream)))
        return out


def lighting(lights, visibility: bool = False, specular_f0=(0.04, 0.04, 0.04)) -> _lib.Lighting:
    """bv_lighting from a list of dicts (position, direction, intensity, exponent,
    and an optional per-light visibility flag); visibility=True turns the
    visibility ray on for every light."""
    if not 1 <= len(lights) <= 8:
        raise ValueError("1..8 point lights")
    lit = _lib.Lighting()
    lit.count = len(lights)
    lit.visibility = 1 if visibility else 0
    lit.specular_f0[:] = [float(x) for x in specular_f0]
    for i, L in enumerate(lights):
        lit.lights[i].position[:] = [float(x) for x in L["position"]]
        lit.lights[i].direction[:] = [float(x) for x in L.get("direction", (0.0, 0.0, 1.0))]
        lit.lights[i].intensity[:] = [float(x) for x in L["intensity"]]
        lit.lights[i].exponent = float(L.get("exponent", 0.0))
        lit.lights[i].visibility = 1 if L.get("visibility") else 0
    return lit


def relight_camera(cloud: RadFoamGpuCloud, camera: Camera, size, lights, visibility: bool = False,
                   specular_f0=(0.04, 0.04, 0.04), settings: RadFoamTraceSettings = RadFoamTraceSettings(),
                   start_cell: Optional[int] = None, out=None, stream=None):
    """Relit image [H*W, 4] of a scene created with materials."""
    torch = _torch()
    w, h = int(size[0]), int(size[1])
    if out is None:
        out = torch.empty((w * h, 4), dtype=torch.float32, device="cuda")
    st = TraceSettings(settings.weight_threshold, settings.max_steps, float(camera.depth), EVAL_SH)
    sc = NO_CELL if start_cell is None else int(start_cell)
    lit = lighting(lights, visibility, specular_f0)
    check(cloud._lib.bv_relight_camera(cloud.handle, C.byref(_camera_c(camera)), w, h, sc, C.byref(st.c()),
                                       C.byref(lit), out.data_ptr(), _stream_ptr(stream)))
    return out


def relight_rays(cloud: RadFoamGpuCloud, origins, dirs, start_cells, lights, visibility: bool = False,
                 specular_f0=(0.04, 0.04, 0.04), settings: TraceSettings = TraceSettings(), start_cell: int = 0,
                 out=None, stream=None):
    torch = _torch()
    n = origins.shape[0]
    if out is None:
        out = torch.empty((n, 4), dtype=torch.float32, device=origins.device)
    rays = _rays(origins, dirs, start_cells, start_cell)
    lit = lighting(lights, visibility, specular_f0)
    check(cloud._lib.bv_relight_rays(cloud.handle, C.byref(rays), C.byref(settings.c()), C.byref(lit),
                                     out.data_ptr(), _stream_ptr(stream)))
    return out


def splat_trace_rays(cloud: RadFoamGpuCloud, origins, dirs, start_cells=None, start_cell: int = 0,
                     settings: TraceSettings = TraceSettings(), outputs=("rgba",), out: Optional[Dict] = None,
                     stream=None) -> Dict:
    """trace_powerfoam_splats (trace.rs:1456-1546) over a batch of device rays:
    every support sphere hit, clipped by its adjacency row, composited in
    (near, cell) order.  start_cells / start_cell only seed last_point."""
    n = origins.shape[0]
    if out is None:
        out = _alloc_outputs(n, outputs, origins.device)
    rays = _rays(origins, dirs, start_cells, start_cell)
    check(cloud._lib.bv_splat_trace_rays(cloud.handle, C.byref(rays), C.byref(settings.c()),
                                         C.byref(_outputs_c(out)), _stream_ptr(stream)))
    return out


def splat_record(cloud: RadFoamGpuCloud, origins, dirs, max_steps: int, depth: float = 10_000.0, out=None,
                 stream=None) -> Dict:
    """record_powerfoam_splats_jacobians (trace.rs:1414-1443) without Jacobians:
    [P, L] rows of cell (int32 holding u32 bits) / near / dt and count[P]."""
    torch = _torch()
    n = origins.shape[0]
    if out is None:
        dev = origins.device
        out = dict(cells=torch.full((n, max_steps), -1, dtype=torch.int32, device=dev),
                   nears=torch.full((n, max_steps), float("nan"), dtype=torch.float32, device=dev),
                   dts=torch.full((n, max_steps), float("nan"), dtype=torch.float32, device=dev),
                   count=torch.zeros(n, dtype=torch.int32, device=dev))
    rays = _rays(origins, dirs, None, 0)
    check(cloud._lib.bv_splat_record(cloud.handle, C.byref(rays), int(max_steps), float(depth),
                                     out["cells"].data_ptr(), out["nears"].data_ptr(), out["dts"].data_ptr(),
                                     out["count"].data_ptr(), _stream_ptr(stream)))
    return out


SPLAT_PATH_FIELDS = ("cells", "next_cells", "previous_cells", "dts", "mask", "dt_grad_previous", "dt_grad_current",
                     "dt_grad_next", "dt_reference_tangent", "status")


def splat_record_jacobians(cloud: RadFoamGpuCloud, origins, dirs, max_steps: int, depth: float = 10_000.0,
                           max_path_dt: float = 50.0, out: Optional[Dict] = None, stream=None) -> Dict:
    """The PowerFoam path recorder with Jacobian streams (PathRecorder for power
    cells, jacobian_mode 1): [P, L] cells / next_cells / previous_cells (int32
    holding u32 bits), dts, mask, dt_grad_previous / current / next [P, L, 4],
    dt_reference_tangent, and status [P] (count | truncated bit).  Buffers are
    zero-filled when allocated here, as the trainer clears them."""
    torch = _torch()
    n = origins.shape[0]
    if out is None:
        dev = origins.device
        z = lambda *s, dt=torch.float32: torch.zeros(s, dtype=dt, device=dev)
        out = dict(cells=z(n, max_steps, dt=torch.int32), next_cells=z(n, max_steps, dt=torch.int32),
                   previous_cells=z(n, max_steps, dt=torch.int32), dts=z(n, max_steps), mask=z(n, max_steps),
                   dt_grad_previous=z(n, max_steps, 4), dt_grad_current=z(n, max_steps, 4),
                   dt_grad_next=z(n, max_steps, 4), dt_reference_tangent=z(n, max_steps),
                   status=z(n, dt=torch.int32))
    rays = _rays(origins, dirs, None, 0)
    st = _lib.RecordSettings(int(max_steps), float(depth), float(max_path_dt), 0)
    oc = _lib.SplatPathOutputs(*[_ptr(out.get(k)) for k in SPLAT_PATH_FIELDS])
    check(cloud._lib.bv_splat_record_jacobians(cloud.handle, C.byref(rays), C.byref(st), C.byref(oc),
                                               _stream_ptr(stream)))
    return out


def splat_relight_camera(cloud: RadFoamGpuCloud, camera: Camera, size, lights, visibility: bool = False,
                         specular_f0=(0.04, 0.04, 0.04), settings: RadFoamTraceSettings = RadFoamTraceSettings(),
                         out=None, stream=None):
    """Relit PowerFoam splat image [H*W, 4] of a scene created with materials."""
    torch = _torch()
    w, h = int(size[0]), int(size[1])
    if out is None:
        out = torch.empty((w * h, 4), dtype=torch.float32, device="cuda")
    st = TraceSettings(settings.weight_threshold, settings.max_steps, float(camera.depth), EVAL_SH)
    lit = lighting(lights, visibility, specular_f0)
    check(cloud._lib.bv_splat_relight_camera(cloud.handle, C.byref(_camera_c(camera)), w, h, C.byref(st.c()),
                                             C.byref(lit), out.data_ptr(), _stream_ptr(stream)))
    return out


def splat_relight_rays(cloud: RadFoamGpuCloud, origins, dirs, lights, visibility: bool = False,
                       specular_f0=(0.04, 0.04, 0.04), settings: TraceSettings = TraceSettings(), out=None,
                       stream=None):
    torch = _torch()
    n = origins.shape[0]
    if out is None:
        out = torch.empty((n, 4), dtype=torch.float32, device=origins.device)
    rays = _rays(origins, dirs, None, 0)
    lit = lighting(lights, visibility, specular_f0)
    check(cloud._lib.bv_splat_relight_rays(cloud.handle, C.byref(rays), C.byref(settings.c()), C.byref(lit),
                                           out.data_ptr(), _stream_ptr(stream)))
    return out


class PowerFoamGpuSplatTracer:
    """gpu/powerfoam_splat.rs:36-330: the compute-splat PowerFoam renderer.  The
    reference gathers candidates per tile and integrates recorded rows; here the
    whole image is one launch over the device sphere BVH, so there is no
    candidate capacity to validate and batched rows equal the full image by
    construction (pixels are independent)."""

    def __init__(self, cloud: RadFoamGpuCloud, settings: RadFoamTraceSettings = RadFoamTraceSettings(),
                 resolution=None):
        self.cloud = cloud
        self.settings = settings
        self.resolution = None if resolution is None else (int(resolution[0]), int(resolution[1]))

    def reconfigure(self, resolution, max_steps: int) -> None:
        self.resolution = (int(resolution[0]), int(resolution[1]))
        self.settings.max_steps = int(max_steps)

    def set_weight_threshold(self, weight_threshold: float) -> None:
        self.settings.weight_threshold = weight_threshold

    def dispatch(self, camera: Camera, size=None, out=None, outputs=("rgba",), eval_mode: int = EVAL_SH,
                 stream=None) -> Dict:
        """Render (width, height) (default: the configured resolution); out['rgba'] is [H*W, 4]."""
        w, h = (int(size[0]), int(size[1])) if size is not None else self.resolution
        if out is None:
            out = _alloc_outputs(w * h, outputs, "cuda")
        st = TraceSettings(self.settings.weight_threshold, self.settings.max_steps, float(camera.depth), eval_mode)
        check(self.cloud._lib.bv_splat_trace_camera(self.cloud.handle, C.byref(_camera_c(camera)), w, h,
                                                    C.byref(st.c()), C.byref(_outputs_c(out)), _stream_ptr(stream)))
        return out


@dataclass
class RadFoamDepthSettings:
    """gpu/radfoam_depth.rs:6-20 (defaults max_steps 1024, weight_threshold 0.001)."""
    max_steps: int = 1024
    weight_threshold: float = 0.001


def _depth_image(out: Dict, w: int, h: int):
    """The depth tracers' rgba32float texture (radfoam_depth.wgsl:29-35,
    powerfoam_depth.wgsl:143-147): per pixel (mode depth, accumulated alpha,
    peak segment weight, 1), as an [H, W, 4] device tensor."""
    torch = _torch()
    pw = out["peak_weight"]
    return torch.stack([out["depth_mode"], out["rgba"][:, 3], pw, torch.ones_like(pw)], 1).view(h, w, 4)


class RadFoamGpuDepthTracer:
    """gpu/radfoam_depth.rs:39-124: the RadFoam / PowerFoam cell walk seeded at
    the camera's containing cell, integrating density only (no SH), recording
    the heaviest segment: depth_mode = the middle of that segment, peak_weight =
    its weight T * alpha (radfoam_trace.wgsl:164-191).  This is the camera walk
    in opacity mode with the depth outputs."""

    def __init__(self, cloud: RadFoamGpuCloud, settings: RadFoamDepthSettings = RadFoamDepthSettings(),
                 kernel: int = KERNEL_FAST):
        self.cloud = cloud
        self.settings = settings
        self.kernel = kernel

    def dispatch(self, camera: Camera, size, start_cell: Optional[int] = None, stream=None):
        """The (width, height) depth map as [H, W, 4]: (depth_mode, alpha, peak_weight, 1)."""
        w, h = int(size[0]), int(size[1])
        if w <= 0 or h <= 0:
            raise ValueError("depth map must be non-empty")
        out = _alloc_outputs(w * h, ("rgba", "depth_mode", "peak_weight"), "cuda")
        st = TraceSettings(self.settings.weight_threshold, self.settings.max_steps, float(camera.depth), EVAL_OPACITY)
        sc = NO_CELL if start_cell is None else int(start_cell)
        check(self.cloud._lib.bv_trace_camera(self.cloud.handle, C.byref(_camera_c(camera)), w, h, sc,
                                              C.byref(st.c()), C.byref(_outputs_c(out)), self.kernel,
                                              _stream_ptr(stream)))
        return _depth_image(out, w, h)


class PowerFoamGpuDepthTracer:
    """gpu/powerfoam_depth.rs:39-210: PowerFoam depth statistics from
    independently clipped supports (every component of a Čech graph, not only
    the camera's): the splat walk over the sphere BVH in opacity mode, keeping
    the heaviest segment (depth_mode = its entry depth + dt / 2,
    powerfoam_depth.wgsl:118-147)."""

    def __init__(self, cloud: RadFoamGpuCloud, settings: RadFoamDepthSettings = RadFoamDepthSettings(),
                 resolution=(1, 1)):
        if not cloud.bounded:
            raise ValueError("PowerFoam depth tracing requires radii")
        w, h = int(resolution[0]), int(resolution[1])
        if w <= 0 or h <= 0:
            raise ValueError("PowerFoam depth resolution must be non-zero")
        self.cloud = cloud
        self.settings = settings
        self.resolution = (w, h)

    def dispatch(self, camera: Camera, stream=None):
        """The configured resolution's depth map as [H, W, 4]: (depth_mode, alpha, peak_weight, 1)."""
        w, h = self.resolution
        out = _alloc_outputs(w * h, ("rgba", "depth_mode", "peak_weight"), "cuda")
        st = TraceSettings(self.settings.weight_threshold, self.settings.max_steps, float(camera.depth), EVAL_OPACITY)
        check(self.cloud._lib.bv_splat_trace_camera(self.cloud.handle, C.byref(_camera_c(camera)), w, h,
                                                    C.byref(st.c()), C.byref(_outputs_c(out)), _stream_ptr(stream)))
        return _depth_image(out, w, h)


@dataclass
class Transform:
    """Transform (scene.rs:109-175): position, rotation quaternion (x, y, z, w)
    and per-axis scale; the object-to-world matrix is scale, then rotation,
    then translation (Mat4::from_scale_rotation_translation)."""
    position: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    scale: tuple = (1.0, 1.0, 1.0)

    def to_matrix(self) -> np.ndarray:
        """4x4 object-to-world; raises ValueError where to_matrix asserts (non-finite
        position or scale, a scale component below 1e-8, a zero rotation)."""
        p = np.asarray(self.position, np.float64)
        sc = np.asarray(self.scale, np.float64)
        q = np.asarray(self.rotation, np.float64)
        if not (np.all(np.isfinite(p)) and np.all(np.isfinite(sc)) and np.min(np.abs(sc)) > 1e-8):
            raise ValueError("transform position and scale must be finite and the scale invertible")
        n2 = float(q @ q)
        if not (np.isfinite(n2) and n2 > 1e-12):
            raise ValueError("transform rotation must be finite and non-zero")
        x, y, z, w = q / np.sqrt(n2)
        r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        m = np.eye(4)
        m[:3, :3] = r * sc[None, :]
        m[:3, 3] = p
        return m

    def world_to_object(self) -> np.ndarray:
        """Row-major 3x4 FP32 inverse (GpuTransform::world_to_object)."""
        return np.linalg.inv(self.to_matrix())[:3, :].astype(np.float32)


def bounding_sphere(xyz, radius=None):
    """compute_bounding_sphere (scene.rs:253-274, 294-304): centroid and the largest
    distance to it plus the site's support radius (0 for plain RadFoam)."""
    p = np.asarray(xyz, np.float64)
    if p.shape[0] == 0:
        return np.zeros(3), 0.0
    c = p.mean(axis=0)
    dist = np.sqrt(((p - c) ** 2).sum(axis=1))
    if radius is not None:
        dist = dist + np.asarray(radius, np.float64)
    return c, float(dist.max())


class Scene:
    """Multi-object scene of RadFoam / PowerFoam clouds (scene.rs:216-570) traced by
    bv_trace_objects (trace_scene, scene_trace_core.wgsl:406-465).  Objects keep a
    local bounding sphere; its world form is the transformed centre and the radius
    times the largest absolute scale (scene.rs:436-447).  update_camera sets each
    object's start cell to the cell containing the camera's object-space position
    (scene.rs:395-413).  Objects of one scene share an SH degree and kind."""

    def __init__(self):
        self._lib = _lib.load()
        self._objects = []

    def object_count(self) -> int:
        return len(self._objects)

    def add_radfoam(self, cloud: RadFoamGpuCloud, transform: Transform = Transform()) -> int:
        if len(self._objects) >= _lib.MAX_OBJECTS:
            raise ValueError(f"a scene holds at most {_lib.MAX_OBJECTS} objects")
        c, r = bounding_sphere(cloud.host.xyz, cloud.host.radius)
        transform.to_matrix()
        self._objects.append(dict(cloud=cloud, transform=transform, center=c, radius=r, start=0))
        return len(self._objects) - 1

    def set_transform(self, handle: int, transform: Transform) -> None:
        transform.to_matrix()
        self._objects[handle]["transform"] = transform

    def get_transform(self, handle: int) -> Transform:
        return self._objects[handle]["transform"]

    def update_camera(self, world_position) -> None:
        w = np.asarray(world_position, np.float64)
        if not np.all(np.isfinite(w)):
            raise ValueError("camera position must be finite")
        for ob in self._objects:
            local = np.linalg.inv(ob["transform"].to_matrix()) @ np.append(w, 1.0)
            ob["start"] = ob["cloud"].containing_point(local[:3].astype(np.float32))

    def bounds(self, handle: int):
        """World bounding sphere (centre, radius) of one object, as bv_trace_objects sees it."""
        ob = self._objects[handle]
        t = ob["transform"]
        c = (t.to_matrix() @ np.append(ob["center"], 1.0))[:3]
        return c.astype(np.float32), np.float32(ob["radius"] * float(np.max(np.abs(np.asarray(t.scale)))))

    def objects_c(self):
        arr = (_lib.SceneObject * max(1, len(self._objects)))()
        for k, ob in enumerate(self._objects):
            c, r = self.bounds(k)
            arr[k].scene = ob["cloud"].handle.value
            arr[k].world_to_object[:] = [float(v) for v in ob["transform"].world_to_object().reshape(12)]
            arr[k].center[:] = [float(v) for v in c]
            arr[k].radius = float(r)
            arr[k].start_cell = int(ob["start"])
        return arr

    def trace_rays(self, origins, dirs, settings: TraceSettings = TraceSettings(), out=None, stream=None):
        """rgba [n, 4] of world rays (device tensors); call update_camera first with
        the rays' common origin."""
        torch = _torch()
        n = origins.shape[0]
        rays = _rays(origins, dirs, None, 0)
        if out is None:
            out = torch.empty((n, 4), dtype=torch.float32, device=origins.device)
        assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and tuple(out.shape) == (n, 4)
        check(self._lib.bv_trace_objects(self.objects_c(), len(self._objects), C.byref(rays), C.byref(settings.c()),
                                         out.data_ptr(), _stream_ptr(stream)))
        return out

    def dispatch(self, camera: Camera, size, settings: TraceSettings = TraceSettings(), stream=None):
        """One frame: pixel-centre camera rays through every object; [H, W, 4] rgba."""
        w, h = int(size[0]), int(size[1])
        self.update_camera(camera.position)
        o, d = camera_rays(camera, w, h)
        return self.trace_rays(o, d, settings, stream=stream).reshape(h, w, 4)


class Server:
    """The serving loop (bv_server_*): resident scenes, admission by result
    memory and continuous batching of view requests through one persistent
    ray-tile kernel fed by the server's scheduler thread."""

    def __init__(self, hbm_budget_bytes: int = 1 << 30, max_active: int = 64, tiles_ahead: int = 0):
        self._lib = _lib.load()
        cfg = _lib.ServerConfig(int(hbm_budget_bytes), int(max_active), int(tiles_ahead))
        h = C.c_void_p()
        check(self._lib.bv_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h
        self._keep = {}
        _live_servers.add(self)

    def add_scene(self, cloud: RadFoamGpuCloud) -> int:
        out = C.c_uint32()
        check(self._lib.bv_server_add_scene(self.handle, cloud.handle, C.byref(out)))
        return int(out.value)

    def submit(self, scene: int, camera: Camera, size, mode: int = EVAL_SH, lights=None, visibility: bool = False,
               specular_f0=(0.04, 0.04, 0.04), settings: RadFoamTraceSettings = RadFoamTraceSettings(),
               start_cell: Optional[int] = None, out: Optional[np.ndarray] = None, renderer: int = RENDER_WALK):
        """Queue a view; returns (ticket, out) where out [H*W, 4] float32 is filled on completion.
        renderer: RENDER_WALK (cell walk) or RENDER_SPLAT (PowerFoam splats)."""
        w, h = int(size[0]), int(size[1])
        if out is None:
            out = np.empty((w * h, 4), np.float32)
        assert out.dtype == np.float32 and out.flags["C_CONTIGUOUS"] and out.size == w * h * 4
        req = _lib.ViewRequest()
        req.scene, req.width, req.height, req.mode = scene, w, h, mode
        req.start_cell = NO_CELL if start_cell is None else int(start_cell)
        req.renderer = int(renderer)
        req.camera = _camera_c(camera)
        req.settings = TraceSettings(settings.weight_threshold, settings.max_steps, float(camera.depth), mode).c()
        if mode == EVAL_RELIGHT:
            req.lighting = lighting(lights, visibility, specular_f0)
        req.host_rgba = out.ctypes.data
        t = C.c_uint64()
        check(self._lib.bv_server_submit(self.handle, C.byref(req), C.byref(t)))
        self._keep[int(t.value)] = out
        return int(t.value), out

    def poll(self, ticket: int) -> bool:
        d = C.c_int()
        check(self._lib.bv_server_poll(self.handle, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int) -> float:
        """Block until the image is in host memory; returns its latency in ms."""
        lat = C.c_double()
        check(self._lib.bv_server_wait(self.handle, ticket, C.byref(lat)))
        self._keep.pop(ticket, None)
        return float(lat.value)

    def stats(self) -> Dict:
        st = _lib.ServerStats()
        check(self._lib.bv_server_get_stats(self.handle, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(self._lib.bv_server_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Servers still open at interpreter exit are closed, so their kernel has
# drained before the process ends.
_live_servers: "weakref.WeakSet[Server]" = weakref.WeakSet()


@atexit.register
def _close_servers() -> None:
    for srv in list(_live_servers):
        try:
            srv.close()
        except Exception:
            pass


@dataclass
class PathRecordBuffers:
    """gpu/path_record.rs:614-670: [P, L] rows plus a status word per path, the
    batched recorder's pixel-index input [P] and, for power cells, the
    previous-cell and Jacobian streams.  ray_origins / ray_dirs / view_index
    ([P, 3], [P, 3], [P]) are optional per-row ray outputs of dispatch_batch."""
    cells: "object"
    next_cells: "object"
    dts: "object"
    mask: "object"
    status: "object"
    pixel_indices: "object" = None
    previous_cells: "object" = None
    dt_grad_previous: "object" = None
    dt_grad_current: "object" = None
    dt_grad_next: "object" = None
    dt_reference_tangent: "object" = None
    ray_origins: "object" = None
    ray_dirs: "object" = None
    view_index: "object" = None

    @property
    def num_pixels(self) -> int:
        return int(self.cells.shape[0])

    @property
    def max_steps(self) -> int:
        return int(self.cells.shape[1])

    def write_pixel_indices(self, indices) -> None:
        """path_record.rs:1194-1201: one pixel index per row."""
        if len(indices) != self.num_pixels:
            raise ValueError("pixel index slice length must equal num_pixels")
        self.write_pixel_indices_prefix(indices)

    def write_pixel_indices_prefix(self, indices) -> None:
        """path_record.rs:1206-1218: the first len(indices) rows; the rest keep their values."""
        torch = _torch()
        if len(indices) > self.num_pixels:
            raise ValueError("pixel index prefix exceeds buffer capacity")
        src = torch.as_tensor(np.asarray(indices, np.int64).astype(np.uint32).view(np.int32))
        self.pixel_indices[:len(indices)].copy_(src)

    def rows_c(self) -> _lib.SplatPathOutputs:
        return _lib.SplatPathOutputs(*[_ptr(getattr(self, k)) for k in SPLAT_PATH_FIELDS])


@dataclass
class RecordPathsArgs:
    """RecordPathsArgs (gpu/path_record.rs:142-161): one camera slice."""
    camera: Camera
    image_width: int
    image_height: int
    num_pixels: int
    max_steps: int
    start_point: int = 0
    pixel_offset: int = 0
    image_pixel_offset: int = 0
    max_path_dt: float = 50.0
    depth: float = 10_000.0

    def c(self) -> _lib.RecordView:
        return _lib.RecordView(_camera_c(self.camera), int(self.start_point), int(self.pixel_offset),
                               int(self.image_pixel_offset), int(self.image_width), int(self.image_height),
                               int(self.num_pixels), float(self.depth), float(self.max_path_dt))


class PathRecorder:
    """gpu/path_record.rs:236-595: Voronoi cells record with the plane walk,
    power cells with the splat recorder and its Jacobian streams."""

    def __init__(self, cloud: RadFoamGpuCloud):
        self.cloud = cloud

    def buffers(self, paths: int, max_steps: int, device="cuda", rays: bool = False) -> PathRecordBuffers:
        torch = _torch()
        z = lambda *s, dt=torch.float32: torch.zeros(s, dtype=dt, device=device)
        b = PathRecordBuffers(z(paths, max_steps, dt=torch.int32), z(paths, max_steps, dt=torch.int32),
                              z(paths, max_steps), z(paths, max_steps), z(paths, dt=torch.int32),
                              pixel_indices=z(paths, dt=torch.int32))
        if self.cloud.bounded:
            b.previous_cells = z(paths, max_steps, dt=torch.int32)
            b.dt_grad_previous = z(paths, max_steps, 4)
            b.dt_grad_current = z(paths, max_steps, 4)
            b.dt_grad_next = z(paths, max_steps, 4)
            b.dt_reference_tangent = z(paths, max_steps)
        if rays:
            b.ray_origins = z(paths, 3)
            b.ray_dirs = z(paths, 3)
            b.view_index = z(paths, dt=torch.int32)
        return b

    def dispatch(self, origins, dirs, start_cells=None, start_cell: int = 0, *, max_steps: int, depth: float,
                 max_path_dt: float = 50.0, buffers: Optional[PathRecordBuffers] = None,
                 stream=None) -> PathRecordBuffers:
        n = origins.shape[0]
        b = buffers if buffers is not None else self.buffers(n, max_steps, origins.device)
        rays = _rays(origins, dirs, start_cells, start_cell)
        st = _lib.RecordSettings(max_steps, depth, max_path_dt, 0)
        check(self.cloud._lib.bv_record_paths(self.cloud.handle, C.byref(rays), C.byref(st), b.cells.data_ptr(),
                                              b.next_cells.data_ptr(), b.dts.data_ptr(), b.mask.data_ptr(),
                                              b.status.data_ptr(), _stream_ptr(stream)))
        return b

    def dispatch_batch(self, buffers: PathRecordBuffers, args, stream=None) -> PathRecordBuffers:
        """PathRecorder::dispatch_batch (path_record.rs:519-594): every slice's
        rays come from its camera and the pixel indices at its rows."""
        args = list(args)
        for a in args:
            if a.max_steps != buffers.max_steps:
                raise ValueError("path dispatch max_steps must match buffer layout")
        views = (_lib.RecordView * max(len(args), 1))(*[a.c() for a in args])
        out = _lib.RecordBatchOutputs(buffers.rows_c(), _ptr(buffers.ray_origins), _ptr(buffers.ray_dirs),
                                      _ptr(buffers.view_index))
        check(self.cloud._lib.bv_record_paths_batch(self.cloud.handle, views, len(args),
                                                    _ptr(buffers.pixel_indices), buffers.num_pixels,
                                                    buffers.max_steps, C.byref(out), _stream_ptr(stream)))
        return buffers


class PathCompactBuffers:
    """gpu/path_compact.rs:101-230: the dense record stream, storage for
    num_pixels * max_steps records, and its device count."""

    def __init__(self, num_pixels: int, max_steps: int, device="cuda"):
        torch = _torch()
        self.num_pixels, self.max_steps = int(num_pixels), int(max_steps)
        cap = self.num_pixels * self.max_steps
        if cap >= 0xFFFFFFFF:
            raise ValueError("num_pixels * max_steps must fit the 32-bit dense slot index")
        self.count = torch.zeros(1, dtype=torch.int32, device=device)
        self.dense_slots = torch.zeros(cap, dtype=torch.int32, device=device)
        self.cells = torch.zeros(cap, dtype=torch.int32, device=device)
        self.pixel_indices = torch.zeros(cap, dtype=torch.int32, device=device)
        self.active = torch.zeros(cap, dtype=torch.float32, device=device)

    def clear(self, initialize_indices: bool = False) -> None:
        """path_compact.rs:185-195: zero the count and the active mask (and the
        index streams when asked) with device fills."""
        self.count.zero_()
        self.active.zero_()
        if initialize_indices:
            self.dense_slots.zero_()
            self.cells.zero_()
            self.pixel_indices.zero_()

    def active_count(self) -> int:
        """path_compact.rs:197-208: the synchronized number of records."""
        return int(self.count.item()) & 0xFFFFFFFF

    def capacity(self) -> int:
        return self.num_pixels * self.max_steps


class PathCompactor:
    """gpu/path_compact.rs:39-95: appends every recorded path's live prefix to a
    PathCompactBuffers stream (caller clears it first, as in the reference).
    Paths land in path order: deterministic, and one of the orders the
    reference's per-path atomic allows."""

    def __init__(self):
        self._lib = _lib.load()

    def dispatch(self, paths: PathRecordBuffers, output: PathCompactBuffers, stream=None) -> None:
        num_pixels, max_steps = paths.cells.shape
        if (num_pixels, max_steps) != (output.num_pixels, output.max_steps):
            raise ValueError("path rows and compact stream disagree on num_pixels / max_steps")
        check(self._lib.bv_compact_paths(paths.status.data_ptr(), paths.cells.data_ptr(), num_pixels, max_steps,
                                         output.count.data_ptr(), output.dense_slots.data_ptr(),
                                         output.cells.data_ptr(), output.pixel_indices.data_ptr(),
                                         output.active.data_ptr(), output.capacity(), _stream_ptr(stream)))


def adjoint(cloud: RadFoamGpuCloud, origins, dirs, start_cells, rgba_forward, d_rgba, grads: Dict,
            settings: TraceSettings = TraceSettings(), start_cell: int = 0, kernel: int = KERNEL_FAST,
            stream=None) -> Dict:
    """Adds dL/d(density, sh, xyz) into grads['density'] [N], grads['sh'] [N,3K], grads['xyz'] [N,3]."""
    rays = _rays(origins, dirs, start_cells, start_cell)
    g = _lib.Grads(grads["density"].data_ptr(), grads["sh"].data_ptr(), grads["xyz"].data_ptr())
    check(cloud._lib.bv_adjoint(cloud.handle, C.byref(rays), C.byref(settings.c()), rgba_forward.data_ptr(),
                                d_rgba.data_ptr(), C.byref(g), kernel, _stream_ptr(stream)))
    return grads


def train_step_l2(cloud: RadFoamGpuCloud, origins, dirs, start_cells, target_rgb, grads: Dict,
                  settings: TraceSettings = TraceSettings(), start_cell: int = 0, rgba_out=None, loss=None,
                  kernel: int = KERNEL_FAST, stream=None):
    """One training batch with an L2 colour loss: forward, dL/d(rgba), adjoint
    (gradients added into grads).  The fast kernel fuses all three per ray."""
    torch = _torch()
    assert target_rgb.is_cuda and target_rgb.dtype == torch.float32 and target_rgb.is_contiguous()
    assert tuple(target_rgb.shape) == (origins.shape[0], 3)
    rays = _rays(origins, dirs, start_cells, start_cell)
    g = _lib.Grads(grads["density"].data_ptr(), grads["sh"].data_ptr(), grads["xyz"].data_ptr())
    check(cloud._lib.bv_train_step_l2(cloud.handle, C.byref(rays), C.byref(settings.c()), target_rgb.data_ptr(),
                                      _ptr(rgba_out), C.byref(g), _ptr(loss), kernel, _stream_ptr(stream)))
    return grads


def train_step_l2_traced(cloud: RadFoamGpuCloud, origins, dirs, start_cells, target_rgb, grads: Dict, path_len: int,
                         settings: TraceSettings = TraceSettings(), start_cell: int = 0, rgba_out=None, loss=None,
                         stream=None):
    """train_step_l2 with the fast kernel, also returning each ray's visited-cell
    sequence from the fused kernel's forward walk: [n, path_len] int32 (u32
    bits, -1 where unwritten)."""
    torch = _torch()
    n = origins.shape[0]
    path = torch.full((n, path_len), -1, dtype=torch.int32, device=origins.device)
    rays = _rays(origins, dirs, start_cells, start_cell)
    g = _lib.Grads(grads["density"].data_ptr(), grads["sh"].data_ptr(), grads["xyz"].data_ptr())
    check(cloud._lib.bv_train_step_l2_traced(cloud.handle, C.byref(rays), C.byref(settings.c()),
                                             target_rgb.data_ptr(), _ptr(rgba_out), C.byref(g), _ptr(loss),
                                             path.data_ptr(), int(path_len), _stream_ptr(stream)))
    return path


def train_step_l2_noalloc(cloud: RadFoamGpuCloud, origins, dirs, start_cells, target_rgb, grads: Dict,
                          settings: TraceSettings = TraceSettings(), start_cell: int = 0, rgba_out=None, loss=None,
                          stream=None):
    """train_step_l2 with the fast kernel, allocating nothing: the library's
    fixed per-GPU workspace holds the ray counter and path cache.  Calls that
    overlap on one GPU share that workspace and must not run concurrently."""
    torch = _torch()
    assert target_rgb.is_cuda and target_rgb.dtype == torch.float32 and target_rgb.is_contiguous()
    assert tuple(target_rgb.shape) == (origins.shape[0], 3)
    rays = _rays(origins, dirs, start_cells, start_cell)
    g = _lib.Grads(grads["density"].data_ptr(), grads["sh"].data_ptr(), grads["xyz"].data_ptr())
    check(cloud._lib.bv_train_step_l2_noalloc(cloud.handle, C.byref(rays), C.byref(settings.c()),
                                              target_rgb.data_ptr(), _ptr(rgba_out), C.byref(g), _ptr(loss),
                                              _stream_ptr(stream)))
    return grads


def zero_grads(cloud: RadFoamGpuCloud, device="cuda") -> Dict:
    torch = _torch()
    return dict(density=torch.zeros(cloud.n_sites, dtype=torch.float32, device=device),
                sh=torch.zeros((cloud.n_sites, 3 * cloud.K), dtype=torch.float32, device=device),
                xyz=torch.zeros((cloud.n_sites, 3), dtype=torch.float32, device=device))


def l2_loss_grad(rgba, target_rgb, d_rgba=None, loss=None, stream=None):
    """d_rgba = 2 (rgb - target), alpha gradient 0; loss[0] += sum of squared errors."""
    torch = _torch()
    n = rgba.shape[0]
    if d_rgba is None:
        d_rgba = torch.empty((n, 4), dtype=torch.float32, device=rgba.device)
    lib = _lib.load()
    check(lib.bv_l2_loss_grad(rgba.data_ptr(), target_rgb.data_ptr(), n, d_rgba.data_ptr(), _ptr(loss),
                              _stream_ptr(stream)))
    return d_rgba


# ---------------------------------------------------------------- trainer graph

GRAPH_DT_POSITIONS, GRAPH_DT_RECORDED, GRAPH_DT_LINEARIZED = 0, 1, 2
LOSS_L1, LOSS_SMOOTH_L1, LOSS_L2 = 0, 1, 2


@dataclass
class VolumeGraphOptions:
    """The graph builder's options that shape one step (build_volumetric_graph_with_options,
    diff_render.rs:943-965): density activation, background, colour loss,
    opacity loss weight, and how dt is formed."""
    softplus_beta: float = 0.0
    background: tuple = (0.0, 0.0, 0.0)
    loss: int = LOSS_L1
    opacity_weight: float = 0.0
    dt_mode: int = GRAPH_DT_POSITIONS

    def c(self) -> _lib.GraphOptions:
        o = _lib.GraphOptions()
        o.softplus_beta = float(self.softplus_beta)
        for k in range(3):
            o.background[k] = float(self.background[k])
        o.loss, o.opacity_weight, o.dt_mode = int(self.loss), float(self.opacity_weight), int(self.dt_mode)
        return o


def volume_graph_step(params: Dict, batch: Dict, options: VolumeGraphOptions = VolumeGraphOptions(),
                      grads: Optional[Dict] = None, outputs=("loss", "pixels", "opacity", "dt"), stream=None) -> Dict:
    """One forward + backward of the trainer's graph (session.step() without
    the optimizer) over recorded [P, L] path rows, on device tensors.

    params: log_density [N], positions [N,3], sh [N,3K], sh_degree, optional
    log_radii [N] and exposure [V,3].  batch: cells, next_cells, dts, mask
    [P,L]; optional status [P]; previous_cells, dt_reference_tangent [P,L],
    dt_grad_previous / _current / _next [P,L,4] for the linearized dt;
    ray_origin, ray_dir [P,3]; optional basis [P,K-1], view_idx [P]; labels
    [P,3]; optional target_alpha [P].  grads (added to): d_log_density,
    d_positions, d_sh, d_log_radii, d_exposure (any may be missing).  Returns
    the requested outputs (loss [1], pixels [P,3], opacity [P], dt [P,L])."""
    torch = _torch()
    lib = _lib.load()
    P, L = batch["cells"].shape
    dev = batch["cells"].device
    n = params["log_density"].shape[0]
    exposure = params.get("exposure")
    pm = _lib.GraphParams(n, int(params["sh_degree"]), 0 if exposure is None else exposure.shape[0],
                          _ptr(params["log_density"]), _ptr(params["positions"]), _ptr(params["sh"]),
                          _ptr(params.get("log_radii")), _ptr(exposure))
    g = lambda k: _ptr(batch.get(k))
    bt = _lib.GraphBatch(P, L, g("cells"), g("next_cells"), g("dts"), g("mask"), g("status"), g("previous_cells"),
                         g("dt_reference_tangent"), g("dt_grad_previous"), g("dt_grad_current"), g("dt_grad_next"),
                         g("ray_origin"), g("ray_dir"), g("basis"), g("view_idx"), g("labels"), g("target_alpha"))
    out = {}
    if "loss" in outputs:
        out["loss"] = torch.zeros(1, dtype=torch.float32, device=dev)
    if "pixels" in outputs:
        out["pixels"] = torch.empty((P, 3), dtype=torch.float32, device=dev)
    if "opacity" in outputs:
        out["opacity"] = torch.empty(P, dtype=torch.float32, device=dev)
    if "dt" in outputs:
        out["dt"] = torch.empty((P, L), dtype=torch.float32, device=dev)
    oc = _lib.GraphOutputs(_ptr(out.get("loss")), _ptr(out.get("pixels")), _ptr(out.get("opacity")),
                           _ptr(out.get("dt")))
    grads = grads or {}
    gc = _lib.GraphGrads(*[_ptr(grads.get(k)) for k in ("d_log_density", "d_positions", "d_sh", "d_log_radii",
                                                         "d_exposure")])
    check(lib.bv_volume_graph_step(C.byref(pm), C.byref(bt), C.byref(options.c()), C.byref(oc), C.byref(gc),
                                   _stream_ptr(stream)))
    return out


class RadFoamRender:
    """Custom autograd op (the counterpart of the trainer's meganeura graph):
    rgba = render(density, sh, xyz); backward runs the HIP adjoint."""

    @staticmethod
    def apply(cloud: RadFoamGpuCloud, density, sh, xyz, origins, dirs, start_cells=None,
              settings: TraceSettings = TraceSettings()):
        return _render_fn()(cloud, settings, origins, dirs, start_cells, density, sh, xyz)


_RENDER_FN = None


def _render_fn():
    global _RENDER_FN
    if _RENDER_FN is not None:
        return _RENDER_FN.apply
    torch = _torch()

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, cloud, settings, origins, dirs, start_cells, density, sh, xyz):
            cloud.update(xyz=xyz.detach().contiguous(), density=density.detach().contiguous(),
                         sh=sh.detach().contiguous())
            out = trace_rays(cloud, origins, dirs, start_cells, settings=settings, kernel=KERNEL_FAST)
            ctx.cloud, ctx.settings = cloud, settings
            ctx.save_for_backward(origins, dirs, start_cells if start_cells is not None else torch.empty(0),
                                  out["rgba"])
            ctx.has_start = start_cells is not None
            return out["rgba"]

        @staticmethod
        def backward(ctx, grad_rgba):
            origins, dirs, start, rgba = ctx.saved_tensors
            grads = zero_grads(ctx.cloud, rgba.device)
            adjoint(ctx.cloud, origins, dirs, start if ctx.has_start else None, rgba, grad_rgba.contiguous(), grads,
                    settings=ctx.settings)
            return None, None, None, None, None, grads["density"], grads["sh"], grads["xyz"]

    _RENDER_FN = _Render
    return _Render.apply


# ---------------------------------------------------------------- Gaussians

@dataclass
class GaussianTraceSettings:
    """GaussianTraceSettings (trace.rs:106-126)."""
    min_opacity: float = 0.01
    min_transmittance: float = 0.01
    t_start: float = 0.0
    t_end: float = 10_000.0

    def c(self) -> _lib.GaussianSettings:
        return _lib.GaussianSettings(self.min_opacity, self.min_transmittance, self.t_start, self.t_end)


class GaussianGpuCloud:
    """GaussianGpuCloud (gpu/gaussian.rs:94-280): the particles on the device and
    the BVH of their support ellipsoids at min_opacity (InitParameters)."""

    def __init__(self, mean, opacity, rotation, scale, sh, sh_degree: int, min_opacity: float = 0.01):
        self._lib = _lib.load()
        keep = [_host(mean, np.float32), _host(opacity, np.float32), _host(rotation, np.float32),
                _host(scale, np.float32), _host(sh, np.float32)]
        n = keep[1].shape[0]
        K = (sh_degree + 1) ** 2
        assert keep[0].shape == (n, 3) and keep[2].shape == (n, 4) and keep[3].shape == (n, 3)
        assert keep[4].shape == (n, 3 * K)
        desc = _lib.GaussiansDesc(n, *[a.ctypes.data for a in keep], int(sh_degree), float(min_opacity))
        h = C.c_void_p()
        check(self._lib.bv_gaussians_create(C.byref(desc), C.byref(h)))
        self.handle = h
        self.n = n
        self.sh_degree = sh_degree
        self.K = K
        self.min_opacity = float(min_opacity)
        cnt, nbytes = C.c_int64(), C.c_int64()
        check(self._lib.bv_gaussians_get_info(h, C.byref(cnt), C.byref(nbytes)))
        self.device_bytes = int(nbytes.value)

    @classmethod
    def from_scene(cls, g, min_opacity: float = 0.01) -> "GaussianGpuCloud":
        """From any object with mean / opacity / rotation / scale / sh / sh_degree."""
        return cls(g.mean, g.opacity, g.rotation, g.scale, g.sh, g.sh_degree, min_opacity)

    def update(self, mean=None, opacity=None, rotation=None, scale=None, sh=None, stream=None) -> None:
        for t in (mean, opacity, rotation, scale, sh):
            if t is not None:
                assert t.is_cuda and t.dtype == _torch().float32 and t.is_contiguous()
        check(self._lib.bv_gaussians_update(self.handle, _ptr(mean), _ptr(opacity), _ptr(rotation), _ptr(scale),
                                            _ptr(sh), _stream_ptr(stream)))

    def deinit(self) -> None:
        if getattr(self, "handle", None):
            check(self._lib.bv_gaussians_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.deinit()
        except Exception:
            pass


def _gauss_out(n, device, hits=True, order_len=0, stats=False):
    torch = _torch()
    out = dict(rgba=torch.empty((n, 4), dtype=torch.float32, device=device))
    if hits:
        out["hits"] = torch.empty(n, dtype=torch.int32, device=device)
    if order_len:
        out["order"] = torch.full((n, order_len), -1, dtype=torch.int32, device=device)
    if stats:
        out["stats"] = torch.empty((n, 2), dtype=torch.int32, device=device)
    return out


def _gauss_out_c(out, order_len):
    return _lib.GaussianOutputs(_ptr(out["rgba"]), _ptr(out.get("hits")), _ptr(out.get("order")), int(order_len), 0,
                                _ptr(out.get("stats")))


def gaussian_trace_rays(cloud: GaussianGpuCloud, origins, dirs, settings: GaussianTraceSettings = GaussianTraceSettings(),
                        order_len: int = 0, stats: bool = False, out=None, stream=None) -> Dict:
    """trace_gaussians (trace.rs:695-773) over a batch of device rays: rgba [n,4],
    hits [n] and, with order_len > 0, the composited particle indices."""
    n = origins.shape[0]
    if out is None:
        out = _gauss_out(n, origins.device, True, order_len, stats)
    rays = _rays(origins, dirs, None, 0)
    check(cloud._lib.bv_gaussians_trace_rays(cloud.handle, C.byref(rays), C.byref(settings.c()),
                                             C.byref(_gauss_out_c(out, order_len)), _stream_ptr(stream)))
    return out


class GaussianTracer:
    """The Gaussian viewer's per-pixel trace (gaussian_trace.wgsl driven per
    fragment): one ray per pixel of a camera, rgba [H*W, 4]."""

    def __init__(self, cloud: GaussianGpuCloud, settings: GaussianTraceSettings = GaussianTraceSettings()):
        self.cloud = cloud
        self.settings = settings

    def dispatch(self, camera: Camera, size, out=None, stats: bool = False, stream=None) -> Dict:
        w, h = int(size[0]), int(size[1])
        if out is None:
            out = _gauss_out(w * h, "cuda", True, 0, stats)
        st = GaussianTraceSettings(self.settings.min_opacity, self.settings.min_transmittance, self.settings.t_start,
                                   min(self.settings.t_end, float(camera.depth)))
        check(self.cloud._lib.bv_gaussians_trace_camera(self.cloud.handle, C.byref(_camera_c(camera)), w, h,
                                                        C.byref(st.c()), C.byref(_gauss_out_c(out, 0)),
                                                        _stream_ptr(stream)))
        return out


def gaussian_zero_grads(cloud: GaussianGpuCloud, device="cuda") -> Dict:
    torch = _torch()
    n = cloud.n
    z = lambda *s: torch.zeros(s, dtype=torch.float32, device=device)
    return dict(mean=z(n, 3), scale=z(n, 3), rotation=z(n, 4), opacity=z(n), sh=z(n, 3 * cloud.K))


def gaussian_adjoint(cloud: GaussianGpuCloud, origins, dirs, rgba_fwd, d_rgba, grads: Dict,
                     settings: GaussianTraceSettings = GaussianTraceSettings(), stream=None) -> Dict:
    """Adds dL/d(parameters) for dL/d(rgba) of gaussian_trace_rays into grads."""
    rays = _rays(origins, dirs, None, 0)
    g = _lib.GaussianGrads(*[_ptr(grads[k]) for k in ("mean", "scale", "rotation", "opacity", "sh")])
    check(cloud._lib.bv_gaussians_adjoint(cloud.handle, C.byref(rays), C.byref(settings.c()), _ptr(rgba_fwd),
                                          _ptr(d_rgba.contiguous()), C.byref(g), _stream_ptr(stream)))
    return grads


def _gaussian_render_fn():
    torch = _torch()

    class _GaussianRender(torch.autograd.Function):
        """rgba = trace_gaussians(cloud with these parameters, rays): the custom
        autograd op over the HIP forward and adjoint (the trainer-side
        counterpart of a meganeura graph node).  The parameters are pushed into
        the device cloud (bv_gaussians_update) before the forward."""

        @staticmethod
        def forward(ctx, cloud, settings, mean, opacity, rotation, scale, sh, origins, dirs):
            cloud.update(mean.detach().contiguous(), opacity.detach().contiguous(), rotation.detach().contiguous(),
                         scale.detach().contiguous(), sh.detach().contiguous())
            out = gaussian_trace_rays(cloud, origins, dirs, settings)
            ctx.cloud, ctx.settings = cloud, settings
            ctx.save_for_backward(origins, dirs, out["rgba"])
            return out["rgba"]

        @staticmethod
        def backward(ctx, d_rgba):
            origins, dirs, rgba = ctx.saved_tensors
            grads = gaussian_zero_grads(ctx.cloud, rgba.device)
            gaussian_adjoint(ctx.cloud, origins, dirs, rgba, d_rgba.contiguous(), grads, ctx.settings)
            return (None, None, grads["mean"], grads["opacity"], grads["rotation"], grads["scale"], grads["sh"],
                    None, None)

    return _GaussianRender


class GaussianRender:
    """GaussianRender.apply(cloud, settings, mean, opacity, rotation, scale, sh, origins, dirs) -> rgba."""
    _fn = None

    @classmethod
    def apply(cls, *args):
        if cls._fn is None:
            cls._fn = _gaussian_render_fn()
        return cls._fn.apply(*args)
