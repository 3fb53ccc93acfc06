"""ctypes binding of libbvhip.so (the C ABI declared in include/bv.h).

The library is built in-tree by ``bv.build.build()``.  Loading fails loudly when
it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "lib", "libbvhip.so")

BV_OK = 0
BV_ERR_INVALID = 1
BV_ERR_UNSUPPORTED = 2
BV_ERR_HIP = 3
BV_ERR_OOM = 4
BV_ERR_CAPACITY = 5
KERNEL_PLAIN = 0
KERNEL_FAST = 1
RENDER_WALK, RENDER_SPLAT = 0, 1
EVAL_OPACITY, EVAL_CONSTANT, EVAL_SH, EVAL_RELIGHT = 0, 1, 2, 3
NO_CELL = 0xFFFFFFFF
PATH_TRUNCATED = 0x80000000

# Every symbol include/bv.h declares.
EXPORTS = (
    "bv_last_error", "bv_version", "bv_device_info", "bv_scene_create", "bv_scene_destroy", "bv_scene_get_info",
    "bv_scene_update", "bv_containing_cell", "bv_trace_rays", "bv_trace_camera", "bv_camera_rays",
    "bv_record_paths", "bv_adjoint", "bv_train_step_l2", "bv_l2_loss_grad", "bv_relight_camera", "bv_relight_rays",
    "bv_server_create", "bv_server_add_scene", "bv_server_submit", "bv_server_poll", "bv_server_wait",
    "bv_server_get_stats", "bv_server_destroy", "bv_splat_trace_rays", "bv_splat_trace_camera", "bv_splat_record",
    "bv_splat_relight_camera", "bv_splat_relight_rays", "bv_gaussians_create", "bv_gaussians_destroy",
    "bv_gaussians_get_info", "bv_gaussians_update", "bv_gaussians_trace_rays", "bv_gaussians_trace_camera",
    "bv_gaussians_adjoint", "bv_compute_cech", "bv_adjacency_get_info", "bv_adjacency_read",
    "bv_adjacency_destroy", "bv_compact_paths", "bv_trace_objects", "bv_train_step_l2_noalloc",
    "bv_volume_graph_step", "bv_splat_record_jacobians", "bv_compute_delaunay",
    "bv_train_step_l2_traced", "bv_record_paths_batch",
)


class BvError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"bv error {code}: {message}")
        self.code = code


class SceneDesc(C.Structure):
    _fields_ = [("n_sites", C.c_int64), ("xyz", C.c_void_p), ("radius", C.c_void_p), ("density", C.c_void_p),
                ("sh", C.c_void_p), ("sh_degree", C.c_int32), ("offsets", C.c_void_p), ("neighbors", C.c_void_p),
                ("flags", C.c_uint32), ("material", C.c_void_p)]


class SceneInfo(C.Structure):
    _fields_ = [("n_sites", C.c_int64), ("n_edges", C.c_int64), ("sh_degree", C.c_int32), ("bounded", C.c_int32),
                ("max_degree", C.c_uint32), ("pad", C.c_uint32), ("device_bytes", C.c_int64)]


class TraceSettings(C.Structure):
    _fields_ = [("weight_threshold", C.c_float), ("max_steps", C.c_uint32), ("depth", C.c_float),
                ("eval_mode", C.c_uint32), ("const_rgb", C.c_float * 3), ("pad", C.c_uint32)]


MAX_OBJECTS = 8


class SceneObject(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("world_to_object", C.c_float * 12), ("center", C.c_float * 3),
                ("radius", C.c_float), ("start_cell", C.c_uint32), ("pad", C.c_uint32)]


class Rays(C.Structure):
    _fields_ = [("origins", C.c_void_p), ("directions", C.c_void_p), ("start_cells", C.c_void_p),
                ("start_cell", C.c_uint32), ("pad", C.c_uint32), ("count", C.c_int64)]


class TraceOutputs(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("steps", C.c_void_p), ("last_point", C.c_void_p), ("t_end", C.c_void_p),
                ("depth_mode", C.c_void_p), ("peak_weight", C.c_void_p), ("peak_point", C.c_void_p),
                ("stats", C.c_void_p), ("path", C.c_void_p), ("path_len", C.c_uint32), ("pad", C.c_uint32)]


class CameraParams(C.Structure):
    """CameraParams (camera.rs:9-18), 48 bytes."""
    _fields_ = [("position", C.c_float * 3), ("depth", C.c_float), ("orientation", C.c_float * 4),
                ("fov", C.c_float * 2), ("principal", C.c_float * 2)]


class PointLight(C.Structure):
    """PointLight (relight.rs:176-181)."""
    _fields_ = [("position", C.c_float * 3), ("exponent", C.c_float), ("direction", C.c_float * 3),
                ("visibility", C.c_uint32), ("intensity", C.c_float * 3), ("pad1", C.c_float)]


class Lighting(C.Structure):
    _fields_ = [("count", C.c_uint32), ("visibility", C.c_uint32), ("specular_f0", C.c_float * 3), ("pad", C.c_float),
                ("lights", PointLight * 8)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_active", C.c_uint32), ("tiles_ahead", C.c_uint32)]


class ViewRequest(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("mode", C.c_uint32),
                ("start_cell", C.c_uint32), ("renderer", C.c_uint32), ("pad", C.c_uint32 * 2),
                ("camera", CameraParams),
                ("settings", TraceSettings), ("lighting", Lighting), ("host_rgba", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("mrays_per_s", C.c_double), ("busy_ms", C.c_double), ("peak_bytes", C.c_int64)]


class RecordSettings(C.Structure):
    _fields_ = [("max_steps", C.c_uint32), ("depth", C.c_float), ("max_path_dt", C.c_float), ("pad", C.c_uint32)]


class GaussiansDesc(C.Structure):
    _fields_ = [("count", C.c_int64), ("means", C.c_void_p), ("opacity", C.c_void_p), ("rotations", C.c_void_p),
                ("scales", C.c_void_p), ("sh", C.c_void_p), ("sh_degree", C.c_int32), ("min_opacity", C.c_float)]


class GaussianSettings(C.Structure):
    _fields_ = [("min_opacity", C.c_float), ("min_transmittance", C.c_float), ("t_start", C.c_float),
                ("t_end", C.c_float)]


class GaussianOutputs(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("hits", C.c_void_p), ("order", C.c_void_p), ("order_len", C.c_uint32),
                ("pad", C.c_uint32), ("stats", C.c_void_p)]


class GaussianGrads(C.Structure):
    _fields_ = [("d_mean", C.c_void_p), ("d_scale", C.c_void_p), ("d_rotation", C.c_void_p),
                ("d_opacity", C.c_void_p), ("d_sh", C.c_void_p)]


class GraphParams(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("sh_degree", C.c_int32), ("n_views", C.c_int32), ("log_density", C.c_void_p),
                ("positions", C.c_void_p), ("sh", C.c_void_p), ("log_radii", C.c_void_p), ("exposure", C.c_void_p)]


class GraphBatch(C.Structure):
    _fields_ = [("num_pixels", C.c_uint32), ("max_steps", C.c_uint32), ("cells", C.c_void_p),
                ("next_cells", C.c_void_p), ("dts", C.c_void_p), ("mask", C.c_void_p), ("status", C.c_void_p),
                ("previous_cells", C.c_void_p), ("dt_reference_tangent", C.c_void_p),
                ("dt_grad_previous", C.c_void_p), ("dt_grad_current", C.c_void_p), ("dt_grad_next", C.c_void_p),
                ("ray_origin", C.c_void_p), ("ray_dir", C.c_void_p), ("basis", C.c_void_p),
                ("view_idx", C.c_void_p), ("labels", C.c_void_p), ("target_alpha", C.c_void_p)]


class GraphOptions(C.Structure):
    _fields_ = [("softplus_beta", C.c_float), ("background", C.c_float * 3), ("loss", C.c_uint32),
                ("opacity_weight", C.c_float), ("dt_mode", C.c_uint32), ("pad", C.c_uint32)]


class GraphOutputs(C.Structure):
    _fields_ = [("loss", C.c_void_p), ("pixels", C.c_void_p), ("opacity", C.c_void_p), ("dt", C.c_void_p)]


class GraphGrads(C.Structure):
    _fields_ = [("d_log_density", C.c_void_p), ("d_positions", C.c_void_p), ("d_sh", C.c_void_p),
                ("d_log_radii", C.c_void_p), ("d_exposure", C.c_void_p)]


class SplatPathOutputs(C.Structure):
    _fields_ = [("cells", C.c_void_p), ("next_cells", C.c_void_p), ("previous_cells", C.c_void_p),
                ("dts", C.c_void_p), ("mask", C.c_void_p), ("dt_grad_previous", C.c_void_p),
                ("dt_grad_current", C.c_void_p), ("dt_grad_next", C.c_void_p), ("dt_reference_tangent", C.c_void_p),
                ("status", C.c_void_p)]


class RecordView(C.Structure):
    """RecordPathsArgs (gpu/path_record.rs:142-161) as bv_record_view, 80 bytes."""
    _fields_ = [("camera", CameraParams), ("start_point", C.c_uint32), ("pixel_offset", C.c_uint32),
                ("image_pixel_offset", C.c_uint32), ("image_width", C.c_uint32), ("image_height", C.c_uint32),
                ("num_pixels", C.c_uint32), ("depth", C.c_float), ("max_path_dt", C.c_float)]


class RecordBatchOutputs(C.Structure):
    _fields_ = [("rows", SplatPathOutputs), ("ray_origins", C.c_void_p), ("ray_dirs", C.c_void_p),
                ("view_index", C.c_void_p)]


class Grads(C.Structure):
    _fields_ = [("d_density", C.c_void_p), ("d_sh", C.c_void_p), ("d_xyz", C.c_void_p)]


_lib = None


def load(path: str = None):
    """Load the native library (raises if it is not built).  BV_LIBRARY names
    another build of it (the host-sanitizer build) to load instead."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("BV_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (HIP library required, no fallback)")
    # torch ships its own libamdhip64 / libhsa-runtime64.  Importing it first makes the
    # library's libamdhip64.so.7 dependency bind to that already-loaded copy, so the process
    # holds ONE HIP runtime.  Loaded the other way round, /opt/rocm's runtime opens the GPU
    # first and torch's second runtime then reports "No HIP GPUs are available".
    import torch  # noqa: F401

    lib = C.CDLL(path)
    vp, i32, u32, i64, f32 = C.c_void_p, C.c_int, C.c_uint32, C.c_int64, C.c_float
    lib.bv_last_error.restype = C.c_char_p
    lib.bv_last_error.argtypes = []
    lib.bv_version.restype = i32
    lib.bv_device_info.argtypes = [vp, vp]
    lib.bv_scene_create.argtypes = [vp, vp]
    lib.bv_scene_destroy.argtypes = [vp]
    lib.bv_scene_get_info.argtypes = [vp, vp]
    lib.bv_scene_update.argtypes = [vp, vp, vp, vp, vp]
    lib.bv_containing_cell.argtypes = [vp, vp, vp, vp]
    lib.bv_compute_cech.argtypes = [vp, vp, i64, vp, vp]
    lib.bv_compute_delaunay.argtypes = [vp, i64, vp]
    lib.bv_adjacency_get_info.argtypes = [vp, vp, vp]
    lib.bv_adjacency_read.argtypes = [vp, vp, vp]
    lib.bv_adjacency_destroy.argtypes = [vp]
    lib.bv_trace_rays.argtypes = [vp, vp, vp, vp, i32, vp]
    lib.bv_trace_objects.argtypes = [vp, u32, vp, vp, vp, vp]
    lib.bv_trace_camera.argtypes = [vp, vp, u32, u32, u32, vp, vp, i32, vp]
    lib.bv_camera_rays.argtypes = [vp, u32, u32, vp, vp, vp]
    lib.bv_record_paths.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.bv_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, i32, vp]
    lib.bv_compact_paths.argtypes = [vp, vp, u32, u32, vp, vp, vp, vp, vp, C.c_uint64, vp]
    lib.bv_l2_loss_grad.argtypes = [vp, vp, i64, vp, vp, vp]
    lib.bv_train_step_l2.argtypes = [vp, vp, vp, vp, vp, vp, vp, i32, vp]
    lib.bv_train_step_l2_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    lib.bv_train_step_l2_traced.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, u32, vp]
    lib.bv_volume_graph_step.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.bv_splat_record_jacobians.argtypes = [vp, vp, vp, vp, vp]
    lib.bv_record_paths_batch.argtypes = [vp, vp, u32, vp, u32, u32, vp, vp]
    lib.bv_relight_camera.argtypes = [vp, vp, u32, u32, u32, vp, vp, vp, vp]
    lib.bv_relight_rays.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.bv_splat_trace_rays.argtypes = [vp, vp, vp, vp, vp]
    lib.bv_splat_trace_camera.argtypes = [vp, vp, u32, u32, vp, vp, vp]
    lib.bv_splat_record.argtypes = [vp, vp, u32, f32, vp, vp, vp, vp, vp]
    lib.bv_splat_relight_camera.argtypes = [vp, vp, u32, u32, vp, vp, vp, vp]
    lib.bv_splat_relight_rays.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.bv_gaussians_create.argtypes = [vp, vp]
    lib.bv_gaussians_destroy.argtypes = [vp]
    lib.bv_gaussians_get_info.argtypes = [vp, vp, vp]
    lib.bv_gaussians_update.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.bv_gaussians_trace_rays.argtypes = [vp, vp, vp, vp, vp]
    lib.bv_gaussians_trace_camera.argtypes = [vp, vp, u32, u32, vp, vp, vp]
    lib.bv_gaussians_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.bv_server_create.argtypes = [vp, vp]
    lib.bv_server_add_scene.argtypes = [vp, vp, vp]
    lib.bv_server_submit.argtypes = [vp, vp, vp]
    lib.bv_server_poll.argtypes = [vp, C.c_uint64, vp]
    lib.bv_server_wait.argtypes = [vp, C.c_uint64, vp]
    lib.bv_server_get_stats.argtypes = [vp, vp]
    lib.bv_server_destroy.argtypes = [vp]
    for name in EXPORTS:
        if name != "bv_last_error":
            getattr(lib, name).restype = i32
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != BV_OK:
        msg = _lib.bv_last_error().decode("utf-8", "replace") if _lib is not None else ""
        raise BvError(code, msg)
