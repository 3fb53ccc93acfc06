"""Path-record compaction (bv_compact_paths, PathCompactor) against the oracle.

Reference: compact_paths (path_compact.wgsl:24-50) and its test
gpu_path_compaction_emits_each_active_dense_record_once (tests/
gpu_path_record.rs:1078-1196).  Bar: the device stream is bit-identical to the
oracle's (paths in path order), the count equals the recorded live steps, the
mask past the count stays 0, appends land after the existing count, and records
past the capacity are not written.  At a 1080p frame of 128-step rows the
stream is checked on the device for the properties path order fixes: slots
strictly increasing, pixel = slot / max_steps, cell = the row's cell, active 1
exactly on the prefix.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from bv import scenes
from conftest import c1_camera

pytestmark = pytest.mark.gpu


def _bv():
    import bv

    return bv


def _torch():
    import torch

    return torch


def _status_rows(rng, n, L, mean=None):
    status = (rng.integers(0, L + 1, n) if mean is None else np.minimum(rng.poisson(mean, n), L)).astype(np.uint32)
    status[rng.random(n) < 0.05] = 0
    status[rng.random(n) < 0.05] |= np.uint32(oracle.TRUNCATED_BIT)
    cells = rng.integers(0, 1 << 24, (n, L)).astype(np.uint32)
    return status, cells


def _paths(status, cells):
    bv = _bv()
    torch = _torch()
    n, L = cells.shape
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()
    return bv.PathRecordBuffers(t(cells), None, None, None, t(status))


def _stream(out, n):
    h = lambda t: t[:n].cpu().numpy()
    return dict(dense_slots=h(out.dense_slots).view(np.uint32), cells=h(out.cells).view(np.uint32),
                pixel_indices=h(out.pixel_indices).view(np.uint32), active=h(out.active))


def _assert_stream(got, ref, label):
    for k in ("dense_slots", "cells", "pixel_indices", "active"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{label}: {k}")


@pytest.mark.parametrize("n,L", [(1, 96), (1023, 96), (5000, 96), (4097, 1)])
def test_compact_matches_oracle(n, L):
    print(f"case compact n={n} L={L}", flush=True)
    bv = _bv()
    status, cells = _status_rows(np.random.default_rng(n + L), n, L)
    out = bv.PathCompactBuffers(n, L)
    out.clear(initialize_indices=True)
    bv.PathCompactor().dispatch(_paths(status, cells), out)
    ref = oracle.compact_paths(status, cells, L)
    assert out.active_count() == ref["count"]
    _assert_stream(_stream(out, ref["count"]), ref, f"n={n} L={L}")
    assert float(out.active[ref["count"]:].abs().max() if ref["count"] < out.capacity() else 0.0) == 0.0


def test_compact_appends_after_existing_count():
    print("case compact append", flush=True)
    bv = _bv()
    n, L = 3000, 64
    status, cells = _status_rows(np.random.default_rng(1), n, L, mean=12)
    paths = _paths(status, cells)
    out = bv.PathCompactBuffers(n, L)
    out.clear()
    c = bv.PathCompactor()
    c.dispatch(paths, out)
    first = out.active_count()
    c.dispatch(paths, out)
    ref = oracle.compact_paths(status, cells, L, base=first)
    assert out.active_count() == ref["count"] == 2 * first
    got = _stream(out, 2 * first)
    for k in ("dense_slots", "cells", "pixel_indices", "active"):
        np.testing.assert_array_equal(got[k][first:], ref[k], err_msg=f"append: {k}")
        np.testing.assert_array_equal(got[k][:first], ref[k], err_msg=f"first call: {k}")


def test_compact_respects_capacity():
    """A stream shorter than the records: the prefix is written, nothing past
    the capacity (a sentinel tail stays), and the count still counts every record."""
    print("case compact capacity", flush=True)
    bv = _bv()
    torch = _torch()
    lib = bv._lib.load()
    n, L = 2000, 32
    status, cells = _status_rows(np.random.default_rng(2), n, L)
    paths = _paths(status, cells)
    ref = oracle.compact_paths(status, cells, L)
    cap = ref["count"] // 3
    guard = 4096
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    streams = [torch.full((cap + guard,), -7, dtype=torch.int32, device="cuda") for _ in range(3)]
    active = torch.full((cap + guard,), -7.0, dtype=torch.float32, device="cuda")
    rc = lib.bv_compact_paths(paths.status.data_ptr(), paths.cells.data_ptr(), n, L, count.data_ptr(),
                              streams[0].data_ptr(), streams[1].data_ptr(), streams[2].data_ptr(),
                              active.data_ptr(), cap, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    assert int(count.item()) == ref["count"]
    for t, k in zip(streams, ("dense_slots", "cells", "pixel_indices")):
        h = t.cpu().numpy()
        np.testing.assert_array_equal(h[:cap].view(np.uint32), ref[k][:cap])
        assert (h[cap:] == -7).all()
    a = active.cpu().numpy()
    assert (a[:cap] == 1.0).all() and (a[cap:] == -7.0).all()


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_compact_unaligned_streams(shift):
    """Output streams that start off a 16-byte boundary take the per-record
    stores (the aligned path writes whole 16-byte quads): same stream, and
    the words before and after it keep their sentinel."""
    print(f"case compact unaligned shift={shift}", flush=True)
    bv = _bv()
    torch = _torch()
    lib = bv._lib.load()
    n, L = 1500, 40
    status, cells = _status_rows(np.random.default_rng(20 + shift), n, L)
    paths = _paths(status, cells)
    ref = oracle.compact_paths(status, cells, L, base=5)
    total = ref["count"]
    count = torch.full((1,), 5, dtype=torch.int32, device="cuda")
    bufs = [torch.full((total + shift + 64,), -7, dtype=torch.int32, device="cuda") for _ in range(3)]
    act = torch.full((total + shift + 64,), -7.0, dtype=torch.float32, device="cuda")
    # stream element k sits at buf[shift + k]; records land at base 5 onward
    ptr = lambda t: t.data_ptr() + 4 * shift
    rc = lib.bv_compact_paths(paths.status.data_ptr(), paths.cells.data_ptr(), n, L, count.data_ptr(), ptr(bufs[0]),
                              ptr(bufs[1]), ptr(bufs[2]), ptr(act), total,
                              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0 and int(count.item()) == total
    for t, k in zip(bufs, ("dense_slots", "cells", "pixel_indices")):
        h = t.cpu().numpy()
        assert (h[:shift + 5] == -7).all() and (h[shift + total:] == -7).all()
        np.testing.assert_array_equal(h[shift + 5:shift + total].view(np.uint32), ref[k], err_msg=k)
    a = act.cpu().numpy()
    assert (a[shift + 5:shift + total] == 1.0).all() and (a[:shift + 5] == -7.0).all()
    assert (a[shift + total:] == -7.0).all()


def test_compact_recorded_camera_paths():
    """The reference test's flow: record a camera's paths (96-step rows), then
    compact them; the stream is the oracle's compaction of the recorded rows,
    and every record's dense slot carries the recorder's mask."""
    print("case compact recorded paths", flush=True)
    bv = _bv()
    torch = _torch()
    sc = scenes.uniform_scene(16_384, seed=0)
    cloud = bv.RadFoamGpuCloud(sc)
    cam = c1_camera()
    o, d = bv.camera_rays(cam, 64, 48)
    L = 96
    rec = bv.PathRecorder(cloud).dispatch(o, d, start_cell=cloud.containing_point(cam.position), max_steps=L,
                                          depth=100.0, max_path_dt=50.0)
    out = bv.PathCompactBuffers(64 * 48, L)
    out.clear(initialize_indices=True)
    bv.PathCompactor().dispatch(rec, out)
    status = rec.status.cpu().numpy().view(np.uint32)
    cells = rec.cells.cpu().numpy().view(np.uint32)
    ref = oracle.compact_paths(status, cells, L)
    assert out.active_count() == ref["count"] > 0
    got = _stream(out, ref["count"])
    _assert_stream(got, ref, "recorded camera paths")
    mask = rec.mask.cpu().numpy().reshape(-1)
    assert (mask[got["dense_slots"]] == 1.0).all()
    assert int((status & 0x7FFFFFFF).max()) > 10
    cloud.deinit()


def test_compact_1080p_properties():
    """1920x1080 paths of 128-step rows (about 40 live steps each): checked on
    the device for what path order fixes."""
    print("case compact 1080p", flush=True)
    bv = _bv()
    torch = _torch()
    n, L = 1920 * 1080, 128
    g = torch.Generator(device="cuda").manual_seed(3)
    steps = torch.clamp(torch.poisson(torch.full((n,), 40.0, device="cuda"), generator=g), max=L).to(torch.int32)
    trunc = torch.rand(n, device="cuda", generator=g) < 0.05
    status = steps | (trunc.to(torch.int32) << 31)
    cells = torch.randint(0, 1 << 24, (n, L), dtype=torch.int32, device="cuda", generator=g)
    paths = bv.PathRecordBuffers(cells, None, None, None, status)
    out = bv.PathCompactBuffers(n, L)
    out.clear()
    bv.PathCompactor().dispatch(paths, out)
    total = int(steps.to(torch.int64).sum().item())
    assert out.active_count() == total
    slots = out.dense_slots[:total].to(torch.int64)
    assert bool((slots[1:] > slots[:-1]).all())
    assert bool((out.pixel_indices[:total].to(torch.int64) == slots // L).all())
    assert bool((out.cells[:total] == cells.view(-1)[slots]).all())
    assert bool((out.active[:total] == 1.0).all()) and bool((out.active[total:] == 0.0).all())
    # each path's block is the first steps[p] slots of its row: as many records
    # as live steps, each at a step below the count
    per_pixel = torch.bincount(out.pixel_indices[:total].to(torch.int64), minlength=n)
    assert bool(((slots % L) < steps.to(torch.int64)[slots // L]).all())
    assert bool((per_pixel == steps.to(torch.int64)).all())
