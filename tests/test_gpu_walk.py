"""HIP forward walk and path recorder against the oracle (parity tests proper).

Bar: the visited cells (steps, last cell, peak cell, recorded cell / next-cell
rows, status words) are bit-exact; colour, transmittance, t_end, depth and
peak weight agree within 1e-5 absolute in FP32 (the oracle uses libm expf, the
device its own exp, so only those may differ in the last bits).
"""
import numpy as np
import pytest

import oracle
from bv import scenes
from conftest import c1_camera
import fixtures

pytestmark = pytest.mark.gpu

TOL = 1e-5
# (kernel id, lanes per ray): the plain port, and the cooperative fast walk with
# 1, 4 and 16 lanes per ray (selected through BV_LANES_PER_RAY).
KERNELS = [(0, None), (1, 1), (1, 4), (1, 16), (1, 32)]
KERNEL_IDS = ["plain", "fast-g1", "fast-g4", "fast-g16", "fast-g32"]


@pytest.fixture(params=KERNELS, ids=KERNEL_IDS)
def kernel(request, monkeypatch):
    k, lanes = request.param
    if lanes is not None:
        monkeypatch.setenv("BV_LANES_PER_RAY", str(lanes))
    return k


def _bv():
    import bv

    return bv


def _torch():
    import torch

    return torch


def dev(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t, view=None):
    a = t.cpu().numpy()
    return a.view(view) if view is not None else a


def compare_trace(got, ref, label):
    np.testing.assert_array_equal(host(got["steps"], np.uint32), ref["steps"], err_msg=f"{label}: steps")
    np.testing.assert_array_equal(host(got["last_point"], np.uint32), ref["last_point"], err_msg=f"{label}: last")
    np.testing.assert_array_equal(host(got["peak_point"], np.uint32), ref["peak_point"], err_msg=f"{label}: peak cell")
    for k in ("rgba", "t_end", "depth_mode", "peak_weight"):
        err = float(np.abs(host(got[k]) - ref[k]).max()) if ref[k].size else 0.0
        assert err <= TOL, f"{label}: {k} differs by {err}"


ALL_OUT = ("rgba", "steps", "last_point", "t_end", "depth_mode", "peak_weight", "peak_point")


def trace_both(scene, cloud, o, d, start, kernel, **kw):
    bv = _bv()
    settings = bv.TraceSettings(weight_threshold=kw.get("weight_threshold", 0.001),
                                max_steps=kw.get("max_steps", 1024), depth=kw.get("depth", 100.0),
                                eval_mode=kw.get("eval_mode", bv.EVAL_SH))
    starts = np.broadcast_to(np.asarray(start, np.uint32), (len(o),)).copy()
    got = bv.trace_rays(cloud, dev(o), dev(d), dev(starts.view(np.int32)), settings=settings, kernel=kernel,
                        outputs=ALL_OUT)
    ref = oracle.trace_rays(oracle.of_host(scene), o, d, starts, weight_threshold=settings.weight_threshold,
                            max_steps=settings.max_steps, depth=settings.depth, eval_mode=settings.eval_mode)
    return got, ref


def fixture_scenes():
    pf = fixtures.branching(sh_degree=3)
    rng = np.random.default_rng(7)
    pf.radius = np.where(np.arange(pf.n) < 64, 0.08, 0.03).astype(np.float32) + \
        rng.uniform(0, 0.01, pf.n).astype(np.float32)
    return {
        "branching_sh3": fixtures.branching(sh_degree=3),
        "branching_sh1": fixtures.branching(sh_degree=1),
        "chain": fixtures.chain(32, 0.1, 0.5, 2, dc=(0.3, -0.2, 0.1)),
        "tetra": fixtures.tetra(),
        "powerfoam_branching": pf,
    }


@pytest.mark.parametrize("name", list(fixture_scenes().keys()))
def test_camera_fixture_parity(name, kernel):
    """tests/radfoam_gpu_vs_cpu.rs:226-431: camera at (0,0,-1) looking +z,
    fov 1 rad, principal (0.08, -0.04), start cell 0, weight threshold 1e-4."""
    print(f"case camera_fixture {name} kernel={kernel}", flush=True)
    bv = _bv()
    torch = _torch()
    sc = fixture_scenes()[name]
    cloud = bv.RadFoamGpuCloud(sc)
    cam = scenes.Camera(np.array([0.0, 0.0, -1.0], np.float32), 100.0, np.array([0, 0, 0, 1], np.float32),
                        np.array([1.0, 1.0], np.float32), np.array([0.08, -0.04], np.float32))
    tracer = bv.RadFoamGpuTracer(cloud, bv.RadFoamTraceSettings(max_steps=256, weight_threshold=1e-4), kernel)
    got = tracer.dispatch(cam, (24, 20), start_cell=0, outputs=ALL_OUT)
    torch.cuda.synchronize()
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 24, 20)
    ref = oracle.trace_rays(oracle.of_host(sc), o, d, 0, weight_threshold=1e-4, max_steps=256, depth=100.0)
    compare_trace(got, ref, name)
    assert ref["rgba"][:, 3].max() > 0.05  # the fixture is actually hit
    cloud.deinit()


def test_camera_rays_bit_exact():
    bv = _bv()
    cam = scenes.looking_at((2.0, 1.5, -3.0), (0.0, 0.5, 0.0), 0.9, 4 / 3, 100.0)
    o, d = bv.camera_rays(cam, 37, 23)
    ro, rd = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 37, 23)
    np.testing.assert_array_equal(host(o), ro)
    np.testing.assert_array_equal(host(d), rd)


def test_c1_camera_render_parity(c1_scene, kernel):
    """Config 1: 16,384 uniform sites, Delaunay adjacency, 128x128 camera at z=3."""
    print(f"case c1_camera kernel={kernel}", flush=True)
    bv = _bv()
    torch = _torch()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    cam = c1_camera()
    start = cloud.containing_point(cam.position)
    assert start == scenes.containing_cell(c1_scene, cam.position)
    tracer = bv.RadFoamGpuTracer(cloud, bv.RadFoamTraceSettings(), kernel)
    got = tracer.dispatch(cam, (128, 128), outputs=ALL_OUT)
    torch.cuda.synchronize()
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 128, 128)
    ref = oracle.trace_rays(oracle.of_host(c1_scene), o, d, start, weight_threshold=0.001, max_steps=1024,
                            depth=100.0)
    compare_trace(got, ref, "c1")
    assert ref["steps"].mean() > 5
    cloud.deinit()


@pytest.mark.parametrize("eval_mode", [0, 1, 2])
def test_c1_random_rays_parity(c1_scene, kernel, eval_mode):
    """Rays from random interior origins in random directions, per-ray start cells."""
    print(f"case c1_random kernel={kernel} eval={eval_mode}", flush=True)
    bv = _bv()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    rng = np.random.default_rng(21)
    n = 3000
    o = rng.uniform(-0.9, 0.9, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)  # not normalised: the walk divides by the length
    starts = np.array([scenes.containing_cell(c1_scene, x) for x in o], np.uint32)
    got, ref = trace_both(c1_scene, cloud, o, d, starts, kernel, eval_mode=eval_mode, depth=4.0)
    compare_trace(got, ref, "c1 random")
    cloud.deinit()


@pytest.mark.parametrize("lanes", [None, "16", "32"], ids=["default", "g16", "g32"])
@pytest.mark.parametrize("eval_mode", [0, 2], ids=["opacity", "sh"])
def test_pooled_walk_parity(c1_scene, lanes, eval_mode, monkeypatch):
    """A batch of rays with their own start cells, larger than the groups the
    device holds at once, goes through the persistent pooled walk (a group
    takes its wave's next ray as soon as its walk ends): 196,608 rays from 48
    origins in random order.  Visited cells bit-exact, colour within 1e-5,
    and the visited-cell sequences of 4,096 of them equal the oracle's."""
    print(f"case pooled walk lanes={lanes} eval={eval_mode}", flush=True)
    if lanes is not None:
        monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    bv = _bv()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    rng = np.random.default_rng(33)
    origins = rng.uniform(-0.9, 0.9, size=(48, 3)).astype(np.float32)
    ostart = np.array([scenes.containing_cell(c1_scene, x) for x in origins], np.uint32)
    n = 196_608
    which = rng.integers(0, 48, n)
    o = np.ascontiguousarray(origins[which])
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = ostart[which]
    got, ref = trace_both(c1_scene, cloud, o, d, starts, 1, eval_mode=eval_mode, depth=4.0)
    compare_trace(got, ref, "pooled walk")
    L = 96
    settings = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=4.0, eval_mode=eval_mode)
    out = bv.trace_rays(cloud, dev(o), dev(d), dev(starts.view(np.int32)), settings=settings, path_len=L)
    k = 4096
    paths = oracle.trace_paths(oracle.of_host(c1_scene), o[:k], d[:k], starts[:k], path_len=L,
                               weight_threshold=0.001, max_steps=1024, depth=4.0, eval_mode=eval_mode)
    got_path = host(out["path"], np.uint32)[:k]
    live = np.arange(L)[None, :] < np.minimum(paths["steps"], L)[:, None].astype(np.int64)
    assert np.array_equal(got_path[live], paths["path"][live]), "pooled walk visited other cells"
    assert (got_path[~live] == 0xFFFFFFFF).all()
    cloud.deinit()


@pytest.mark.parametrize("lanes", [None, "4"], ids=["default", "g4"])
def test_record_paths_large_batch(c1_scene, lanes, monkeypatch):
    """bv_record_paths over 196,608 rays from 48 origins in random order (many
    rows truncated at 24 steps): rows and status words bit-exact, slots past
    each row's count untouched."""
    print(f"case large recorder batch lanes={lanes}", flush=True)
    if lanes is not None:
        monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    bv = _bv()
    torch = _torch()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    rng = np.random.default_rng(35)
    origins = rng.uniform(-0.9, 0.9, size=(48, 3)).astype(np.float32)
    ostart = np.array([scenes.containing_cell(c1_scene, x) for x in origins], np.uint32)
    n, L = 196_608, 24
    which = rng.integers(0, 48, n)
    o = np.ascontiguousarray(origins[which])
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = ostart[which]
    rec = bv.PathRecorder(cloud)
    bufs = rec.buffers(n, L)
    for k in ("cells", "next_cells"):
        getattr(bufs, k).fill_(0x5A5A5A5A)
    rec.dispatch(dev(o), dev(d), dev(starts.view(np.int32)), max_steps=L, depth=3.0, max_path_dt=0.5, buffers=bufs)
    torch.cuda.synchronize()
    ref = oracle.record_paths(oracle.of_host(c1_scene), o, d, starts, max_steps=L, depth=3.0, max_path_dt=0.5)
    st = host(bufs.status, np.uint32)
    np.testing.assert_array_equal(st, ref["status"])
    live = np.arange(L)[None, :] < (st & 0x7FFFFFFF)[:, None].astype(np.int64)
    for k in ("cells", "next_cells"):
        got = host(getattr(bufs, k), np.uint32)
        np.testing.assert_array_equal(got[live], ref[k][live], err_msg=k)
        assert np.all(got[~live] == 0x5A5A5A5A), f"{k}: a slot past a row's count was written"
    np.testing.assert_array_equal(host(bufs.dts)[live], ref["dts"][live])
    assert (st >> 31).sum() > 0 and live.sum() > n
    cloud.deinit()


def test_edge_cases(c1_scene, kernel):
    print(f"case edge kernel={kernel}", flush=True)
    bv = _bv()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    o = np.array([[0, 0, 0], [0.1, 0.2, 0.3], [0.5, 0.5, 0.5], [0, 0, 0]], np.float32)
    d = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [np.nan, 0, 1]], np.float32)  # zero and NaN directions
    starts = np.array([scenes.containing_cell(c1_scene, x) for x in o], np.uint32)
    for max_steps in (1, 2, 7):
        got, ref = trace_both(c1_scene, cloud, o, d, starts, kernel, max_steps=max_steps)
        compare_trace(got, ref, f"edge max_steps={max_steps}")
    got, ref = trace_both(c1_scene, cloud, o, d, starts, kernel, depth=1e-3)
    compare_trace(got, ref, "edge tiny depth")
    got, ref = trace_both(c1_scene, cloud, o, d, starts, kernel, weight_threshold=1.0)
    compare_trace(got, ref, "edge threshold 1")
    # an empty batch is a no-op
    empty = bv.trace_rays(cloud, dev(np.zeros((0, 3), np.float32)), dev(np.zeros((0, 3), np.float32)))
    assert empty["rgba"].shape == (0, 4)
    cloud.deinit()


@pytest.mark.parametrize("lanes", [1, 4, 16, 32])
def test_record_paths_parity(c1_scene, lanes, monkeypatch):
    """radfoam_record_paths.wgsl semantics: rows bit-exact, status with truncation bit."""
    print(f"case record_paths lanes={lanes}", flush=True)
    monkeypatch.setenv("BV_LANES_PER_RAY", str(lanes))
    bv = _bv()
    torch = _torch()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    rng = np.random.default_rng(8)
    n = 2048
    o = rng.uniform(-0.9, 0.9, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = np.array([scenes.containing_cell(c1_scene, x) for x in o], np.uint32)
    for L, max_dt in ((160, 50.0), (12, 0.05)):
        rec = bv.PathRecorder(cloud).dispatch(dev(o), dev(d), dev(starts.view(np.int32)), max_steps=L, depth=4.0,
                                              max_path_dt=max_dt)
        torch.cuda.synchronize()
        ref = oracle.record_paths(oracle.of_host(c1_scene), o, d, starts, max_steps=L, depth=4.0, max_path_dt=max_dt)
        status = host(rec.status, np.uint32)
        np.testing.assert_array_equal(status, ref["status"])
        cnt = status & 0x7FFFFFFF
        valid = np.arange(L)[None, :] < cnt[:, None]
        np.testing.assert_array_equal(host(rec.cells, np.uint32)[valid], ref["cells"][valid])
        np.testing.assert_array_equal(host(rec.next_cells, np.uint32)[valid], ref["next_cells"][valid])
        np.testing.assert_array_equal(host(rec.dts)[valid], ref["dts"][valid])
        np.testing.assert_array_equal(host(rec.mask)[valid], 1.0)
        assert np.all(host(rec.mask)[~valid] == 0.0)  # untouched slots keep their pre-dispatch value
        if L == 12:
            assert np.any(status & 0x80000000)
    cloud.deinit()


def test_containing_cell_gpu(c1_scene):
    bv = _bv()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    rng = np.random.default_rng(9)
    for q in rng.uniform(-3, 3, size=(16, 3)).astype(np.float32):
        assert cloud.containing_point(q) == scenes.containing_cell(c1_scene, q)
    cloud.deinit()


def test_containing_cell_gpu_power_cells():
    """The power-nearest site (|x - p|^2 - r^2, lib.rs:526-541) on the device
    for a scene with radii, against the host's FP32 argmin, and an exact tie
    (two sites at the same power distance) going to the lower index."""
    bv = _bv()
    pw = scenes.power_scene(4000, seed=11)
    cloud = bv.RadFoamGpuCloud(pw)
    rng = np.random.default_rng(12)
    for q in rng.uniform(-1.3, 1.3, size=(24, 3)).astype(np.float32):
        assert cloud.containing_point(q) == scenes.containing_cell(pw, q)
    # radii change the answer: the nearest site is not the power-nearest one
    plain = scenes.HostScene(pw.xyz, pw.density, pw.sh, pw.sh_degree, pw.offsets, pw.neighbors)
    differ = sum(scenes.containing_cell(plain, q) != scenes.containing_cell(pw, q)
                 for q in rng.uniform(-1.0, 1.0, size=(64, 3)).astype(np.float32))
    assert differ > 0
    cloud.deinit()
    xyz = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 3.0, 0.0]], np.float32)
    tie = scenes.HostScene(xyz, np.ones(3, np.float32), np.zeros((3, 3), np.float32), 0,
                           np.array([0, 2, 4, 6], np.uint32), np.array([1, 2, 0, 2, 0, 1], np.uint32),
                           np.full(3, 0.25, np.float32))
    c = bv.RadFoamGpuCloud(tie)
    assert c.containing_point(np.zeros(3, np.float32)) == 0
    c.deinit()


def test_scene_update_matches_new_parameters(small_scene):
    """bv_scene_update rewrites every derived layout: rendering after an update
    equals rendering a scene created with the new values."""
    bv = _bv()
    torch = _torch()
    cloud = bv.RadFoamGpuCloud(small_scene)
    rng = np.random.default_rng(10)
    new = scenes.HostScene((small_scene.xyz + rng.normal(scale=1e-3, size=small_scene.xyz.shape)).astype(np.float32),
                           (small_scene.density * 1.5).astype(np.float32),
                           (small_scene.sh + 0.1).astype(np.float32), 3, small_scene.offsets, small_scene.neighbors)
    cloud.update(xyz=dev(new.xyz), density=dev(new.density), sh=dev(new.sh))
    n = 512
    o = rng.uniform(-0.5, 0.5, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = np.array([scenes.containing_cell(new, x) for x in o], np.uint32)
    import os

    for kernel, lanes in KERNELS:
        if lanes is not None:
            os.environ["BV_LANES_PER_RAY"] = str(lanes)
        got, ref = trace_both(new, cloud, o, d, starts, kernel, depth=3.0)
        torch.cuda.synchronize()
        compare_trace(got, ref, f"updated kernel={kernel} lanes={lanes}")
    os.environ.pop("BV_LANES_PER_RAY", None)
    cloud.deinit()
