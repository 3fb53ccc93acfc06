"""The C ABI library loads, exports every symbol include/bv.h declares, keeps the
reference's uniform layouts, and rejects invalid input before touching a GPU.
CPU only: no call here launches a kernel or allocates device memory."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from bv import _lib
from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bv.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(bv_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_layouts_match_reference_sizes():
    # CameraParams is 48 bytes (camera.rs:9-18, the TraceParams/Camera uniform of gpu/radfoam_trace.rs:183-187).
    assert C.sizeof(_lib.CameraParams) == 48
    assert C.sizeof(_lib.TraceSettings) == 32
    assert C.sizeof(_lib.Rays) == 40
    assert C.sizeof(_lib.RecordSettings) == 16


def _desc(n=2, flags=0, offsets=None, nbrs=None, sh_degree=0):
    xyz = np.zeros((n, 3), np.float32)
    dens = np.ones(n, np.float32)
    sh = np.zeros((n, 3), np.float32)
    off = np.array(offsets if offsets is not None else [0] * (n + 1), np.uint32)
    nb = np.array(nbrs if nbrs is not None else [], np.uint32)
    keep = (xyz, dens, sh, off, nb)
    d = _lib.SceneDesc(n, xyz.ctypes.data, None, dens.ctypes.data, sh.ctypes.data, sh_degree, off.ctypes.data,
                       nb.ctypes.data if nb.size else None, flags)
    return d, keep


@pytest.mark.parametrize("case,code", [
    (dict(flags=1), _lib.BV_ERR_UNSUPPORTED),       # oriented PowerFoam is out of scope
    (dict(n=0), _lib.BV_ERR_INVALID),               # "model has no points"
    (dict(sh_degree=4), _lib.BV_ERR_INVALID),
    (dict(offsets=[1, 1, 1]), _lib.BV_ERR_INVALID),  # offsets[0] != 0
    (dict(offsets=[0, 2, 1], nbrs=[1, 0]), _lib.BV_ERR_INVALID),  # decreasing
    (dict(offsets=[0, 1, 2], nbrs=[1, 5]), _lib.BV_ERR_INVALID),  # out of range
    (dict(offsets=[0, 1, 2], nbrs=[0, 0]), _lib.BV_ERR_INVALID),  # self neighbour
])
def test_scene_validation_errors(case, code):
    lib = _lib.load()
    d, keep = _desc(**case)
    h = C.c_void_p()
    rc = lib.bv_scene_create(C.byref(d), C.byref(h))
    assert rc == code
    assert h.value is None
    assert lib.bv_last_error()


def test_null_arguments_are_errors():
    lib = _lib.load()
    assert lib.bv_scene_create(None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_trace_rays(None, None, None, None, 1, None) == _lib.BV_ERR_INVALID
    assert lib.bv_scene_destroy(None) == _lib.BV_OK
    assert lib.bv_l2_loss_grad(None, None, -1, None, None, None) == _lib.BV_ERR_INVALID
    # entries added for the trainer path check their handles before any launch
    assert lib.bv_record_paths_batch(None, None, 0, None, 0, 8, None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_volume_graph_step(None, None, None, None, None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_splat_record_jacobians(None, None, None, None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_compute_delaunay(None, 8, None) == _lib.BV_ERR_INVALID


def test_cech_argument_errors():
    # compute_cech asserts radii.len() == points.len() (adjacency.rs:434-440); the
    # C ABI checks what it can see: counts, null arrays and a null result slot
    lib = _lib.load()
    xyz = np.zeros((2, 3), np.float32)
    r = np.ones(2, np.float32)
    h = C.c_void_p()
    assert lib.bv_compute_cech(xyz.ctypes.data, r.ctypes.data, 2, None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_compute_cech(xyz.ctypes.data, r.ctypes.data, -1, None, C.byref(h)) == _lib.BV_ERR_INVALID
    assert lib.bv_compute_cech(None, r.ctypes.data, 2, None, C.byref(h)) == _lib.BV_ERR_INVALID
    assert lib.bv_compute_cech(xyz.ctypes.data, None, 2, None, C.byref(h)) == _lib.BV_ERR_INVALID
    assert h.value is None
    assert lib.bv_adjacency_get_info(None, None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_adjacency_read(None, None, None) == _lib.BV_ERR_INVALID
    assert lib.bv_adjacency_destroy(None) == _lib.BV_OK
    import bv

    with pytest.raises(ValueError):
        bv.compute_cech(xyz, r[:1])


def test_compact_argument_errors():
    # PathCompactor::dispatch asserts matching buffer shapes (path_compact.rs:69-70);
    # the C ABI checks what it can see before any launch: the count slot, the
    # 32-bit dense index range, null rows and streams, the capacity bound
    lib = _lib.load()
    a = np.zeros(4, np.uint32)
    f = np.zeros(4, np.float32)
    p = a.ctypes.data
    assert lib.bv_compact_paths(p, p, 1, 4, None, p, p, p, f.ctypes.data, 4, None) == _lib.BV_ERR_INVALID
    assert lib.bv_compact_paths(p, p, 1 << 20, 1 << 13, p, p, p, p, f.ctypes.data, 4, None) == _lib.BV_ERR_INVALID
    assert lib.bv_compact_paths(None, None, 0, 4, p, None, None, None, None, 0, None) == _lib.BV_OK
    assert lib.bv_compact_paths(None, p, 1, 4, p, p, p, p, f.ctypes.data, 4, None) == _lib.BV_ERR_INVALID
    assert lib.bv_compact_paths(p, p, 1, 4, p, p, p, p, f.ctypes.data, 0xFFFFFFFF, None) == _lib.BV_ERR_INVALID
    assert lib.bv_compact_paths(p, p, 1, 4, p, None, p, p, f.ctypes.data, 4, None) == _lib.BV_ERR_INVALID
    assert b"capacity" in lib.bv_last_error() or b"null" in lib.bv_last_error()
