"""The benchmark configurations at full size, checked against the oracle.

* C2: the 1,048,576-cell mixture-on-shell scene (SH 3, Delaunay adjacency,
  seed 1); 4,096 rays, 1,024 random pixels of each of four of the 64 ring
  cameras (1024x768, radius 3).  The fast walk (4 and 16 lanes per ray) and
  the plain port: the whole visited-cell sequence of every ray, step counts,
  last and peak cells bit-exact; colour, t_end, depth and peak weight within
  1e-5.  The recorder's [P, 256] rows and status bit-exact.
* C3: 4,096 random (camera, pixel) training rays on the same scene, L2 loss
  against the perturbed scene's colours.  The fused training kernel (its own
  forward walk's cell sequences bit-exact, colour within 1e-5), the plain
  pipeline and the standalone adjoint: gradients within 1e-4 of each table's
  largest entry against the FP64 oracle adjoint; the FP64 oracle adjoint
  against FP64 central differences of the loss on 64 sampled parameters at a
  pure 1e-4 relative bar (the count that needed the absolute floor is
  printed and bounded).
* C4: 524,288 Gaussians (k-buffer 16), 256 rays of the 1024x1024 camera:
  composited particle order bit-exact, colour within 1e-5, adjoint within 1e-4
  of the FP32 oracle and, against FP64, no further than the FP32 restatement
  itself (whose formula loses up to ~5e-4 of a table's maximum on thin
  particles at this size).
* C5: the 2,097,152-site PowerFoam serving scene with its Cech adjacency from
  the device builder; 256 rays of a 1920x1080 view: recorded intervals and the
  Jacobian streams' rows bit-exact, traced colour within 1e-5.
"""
import math

import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def _host(t, u32=False):
    a = t.cpu().numpy()
    return a.view(np.uint32) if u32 else a


@pytest.fixture(scope="module")
def shell():
    import bv

    print("case C2/C3 scene: 1,048,576 shell sites, Delaunay", flush=True)
    scene = scenes.shell_scene(1 << 20, seed=1, sh_degree=3)
    cloud = bv.RadFoamGpuCloud(scene)
    cams = scenes.ring_cameras(64, 3.0, math.radians(60.0), 1024 / 768, 100.0)
    starts = [cloud.containing_point(c.position) for c in cams]
    yield scene, cloud, cams, starts
    cloud.deinit()


def _camera_batch(cams, starts, which, per_cam, seed, W=1024, H=768):
    import bv

    rng = np.random.default_rng(seed)
    O, D, S = [], [], []
    for k in which:
        o, d = bv.camera_rays(cams[k], W, H)
        pix = rng.integers(0, W * H, per_cam)
        O.append(_host(o)[pix])
        D.append(_host(d)[pix])
        S.append(np.full(per_cam, starts[k], np.uint32))
    return np.concatenate(O), np.concatenate(D), np.concatenate(S)


@pytest.mark.parametrize("kernel,lanes", [(1, "4"), (1, "16"), (0, None)], ids=["fast-g4", "fast-g16", "plain"])
def test_c2_walk_sequences(shell, kernel, lanes, monkeypatch):
    import bv

    print(f"case C2 walk kernel={kernel} lanes={lanes}", flush=True)
    if lanes:
        monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    scene, cloud, cams, starts = shell
    o, d, s = _camera_batch(cams, starts, (0, 16, 32, 48), 1024, seed=21)
    st = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=100.0)
    L = 512
    got = bv.trace_rays(cloud, _dev(o), _dev(d), _dev(s), settings=st, kernel=kernel,
                        outputs=("rgba", "steps", "last_point", "t_end", "depth_mode", "peak_weight", "peak_point"),
                        path_len=L)
    osc = oracle.of_host(scene)
    ref = oracle.trace_rays(osc, o, d, s, weight_threshold=0.001, max_steps=1024, depth=100.0)
    paths = oracle.trace_paths(osc, o, d, s, path_len=L, weight_threshold=0.001, max_steps=1024, depth=100.0)
    steps = _host(got["steps"], True)
    assert np.array_equal(steps, ref["steps"]) and np.array_equal(steps, paths["steps"])
    assert steps.max() <= L and steps.mean() > 20
    live = np.arange(L)[None, :] < steps[:, None].astype(np.int64)
    assert np.array_equal(_host(got["path"], True)[live], paths["path"][live]), "visited cells differ"
    for k in ("last_point", "peak_point"):
        assert np.array_equal(_host(got[k], True), ref[k]), k
    for k in ("rgba", "t_end", "depth_mode", "peak_weight"):
        err = float(np.abs(_host(got[k]) - ref[k]).max())
        assert err <= TOL, f"{k} differs by {err}"


def test_c2_recorder_rows(shell):
    import bv

    print("case C2 recorder", flush=True)
    scene, cloud, cams, starts = shell
    o, d, s = _camera_batch(cams, starts, (8, 40), 2048, seed=22)
    L = 256
    b = bv.PathRecorder(cloud).dispatch(_dev(o), _dev(d), _dev(s), max_steps=L, depth=100.0, max_path_dt=50.0)
    ref = oracle.record_paths(oracle.of_host(scene), o, d, s, max_steps=L, depth=100.0, max_path_dt=50.0)
    st = _host(b.status, True)
    assert np.array_equal(st, ref["status"])
    live = np.arange(L)[None, :] < (st & 0x7FFFFFFF)[:, None].astype(np.int64)
    assert np.array_equal(_host(b.cells, True)[live], ref["cells"][live])
    assert np.array_equal(_host(b.next_cells, True)[live], ref["next_cells"][live])
    assert np.array_equal(_host(b.dts)[live], ref["dts"][live])


@pytest.fixture(scope="module")
def c3_batch(shell):
    import bv

    scene, cloud, cams, starts = shell
    rng = np.random.default_rng(23)
    n = 4096
    cam, px, py = scenes.random_rays(cams, 1024, 768, n, rng)
    O = np.zeros((n, 3), np.float32)
    D = np.zeros((n, 3), np.float32)
    for k in np.unique(cam):
        o, d = bv.camera_rays(cams[k], 1024, 768)
        sel = np.nonzero(cam == k)[0]
        pix = py[sel].astype(np.int64) * 1024 + px[sel]
        O[sel] = _host(o)[pix]
        D[sel] = _host(d)[pix]
    S = np.asarray(starts, np.uint32)[cam]
    target_scene = scenes.perturbed(scene, seed=2)
    tref = oracle.trace_rays(oracle.of_host(target_scene), O, D, S, weight_threshold=0.001, max_steps=1024,
                             depth=100.0)
    return O, D, S, np.ascontiguousarray(tref["rgba"][:, :3])


def _l2_loss64(s64, o, d, s, target):
    r = oracle.trace_rays(s64, o, d, s, weight_threshold=0.001, max_steps=1024, depth=100.0, dtype=np.float64)
    e = r["rgba"][:, :3] - target
    return float((e * e).sum()), r


def test_c3_fd_pins_the_fp64_adjoint(shell, c3_batch):
    """64 sampled parameters (density, SH, position) of the cells the batch
    crosses: FP64 central differences (Richardson, steps h and h/2) of the FP64
    loss against the FP64 oracle adjoint, pure relative 1e-4.  A parameter
    whose two quotients disagree straddles a kink (a face switch or the
    transmittance early-out inside the step) and is skipped; the count is
    bounded."""
    print("case C3 FD of the FP64 adjoint", flush=True)
    scene = shell[0]
    o, d, s, target = c3_batch
    sel = np.arange(0, len(o), 16)  # 256 rays keep 3 x 64 x 4 FP64 forwards quick
    o, d, s, target = o[sel], d[sel], s[sel], target[sel].astype(np.float64)
    s64 = oracle.of_host(scene).as_dtype(np.float64)
    base, r = _l2_loss64(s64, o, d, s, target)
    g64 = np.concatenate([2.0 * (r["rgba"][:, :3] - target), np.zeros((len(o), 1))], 1)
    adj = oracle.adjoint(s64, o, d, s, g64, weight_threshold=0.001, depth=100.0, dtype=np.float64)
    rng = np.random.default_rng(24)
    picks = []
    for name, arr, k in (("density", s64.density, 24), ("sh", s64.sh, 24), ("xyz", s64.xyz, 16)):
        g = adj[name].reshape(-1)
        nz = np.nonzero(np.abs(g) > 1e-3 * np.abs(g).max())[0]
        picks += [(name, arr, i, g[i]) for i in rng.choice(nz, size=k, replace=False)]
    checked = kinks = floor = 0
    for name, arr, i, g in picks:
        flat = arr.reshape(-1)
        h = 1e-6 * max(1.0, abs(flat[i]))
        q = []
        for step in (h, h / 2):
            v0 = flat[i]
            flat[i] = v0 + step
            lp = _l2_loss64(s64, o, d, s, target)[0]
            flat[i] = v0 - step
            lm = _l2_loss64(s64, o, d, s, target)[0]
            flat[i] = v0
            q.append((lp - lm) / (2 * step))
        if abs(q[0] - q[1]) > 1e-3 * max(abs(q[1]), 1e-12):
            kinks += 1
            continue
        fd = (4 * q[1] - q[0]) / 3
        rel = abs(g - fd) / max(abs(fd), abs(g))
        if rel > 1e-4:
            floor += 1
            assert abs(g - fd) <= 1e-4 * abs(adj[name]).max(), (name, i, g, fd)
        checked += 1
    print(f"FD: {checked} checked, {floor} needed the floor, {kinks} straddled a kink", flush=True)
    assert checked >= 56 and floor <= 2


def _gpu_grads(cloud):
    import bv

    return bv.zero_grads(cloud)


def _oracle_adjoint(scene, o, d, s, rgba, target):
    g64 = np.concatenate([2.0 * (rgba[:, :3].astype(np.float64) - target), np.zeros((len(o), 1))], 1)
    return oracle.adjoint(oracle.of_host(scene).as_dtype(np.float64), o, d, s, g64, weight_threshold=0.001,
                          depth=100.0, dtype=np.float64)


def _check_grads(grads, ref, label):
    for k in ("density", "sh", "xyz"):
        g = _host(grads[k]).reshape(ref[k].shape)
        scale = float(np.abs(ref[k]).max())
        err = float(np.abs(g - ref[k]).max())
        assert err <= 1e-4 * scale, f"{label}: d_{k} err {err} vs max {scale}"


@pytest.mark.parametrize("kernel", [1, 0], ids=["fused", "plain"])
def test_c3_training_step(shell, c3_batch, kernel):
    import torch

    import bv

    print(f"case C3 training kernel={kernel}", flush=True)
    scene, cloud, _, _ = shell
    o, d, s, target = c3_batch
    st = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=100.0)
    grads = _gpu_grads(cloud)
    rgba = torch.empty((len(o), 4), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    T = _dev(target.astype(np.float32))
    ref_fwd = oracle.trace_rays(oracle.of_host(scene), o, d, s, weight_threshold=0.001, max_steps=1024,
                                depth=100.0)
    L = 512
    if kernel == 1:
        path = bv.train_step_l2_traced(cloud, _dev(o), _dev(d), _dev(s), T, grads, L, settings=st, rgba_out=rgba,
                                       loss=loss)
        paths = oracle.trace_paths(oracle.of_host(scene), o, d, s, path_len=L, weight_threshold=0.001,
                                   max_steps=1024, depth=100.0)
        live = np.arange(L)[None, :] < paths["steps"][:, None].astype(np.int64)
        assert np.array_equal(_host(path, True)[live], paths["path"][live]), "fused forward visited other cells"
    else:
        bv.train_step_l2(cloud, _dev(o), _dev(d), _dev(s), T, grads, settings=st, rgba_out=rgba, loss=loss,
                         kernel=kernel)
    torch.cuda.synchronize()
    got_rgba = _host(rgba)
    err = float(np.abs(got_rgba - ref_fwd["rgba"]).max())
    assert err <= TOL, f"colour differs by {err}"
    e = ref_fwd["rgba"][:, :3].astype(np.float64) - target
    assert abs(float(loss.item()) - float((e * e).sum())) <= 1e-4 * max(1.0, float((e * e).sum()))
    _check_grads(grads, _oracle_adjoint(scene, o, d, s, ref_fwd["rgba"], target.astype(np.float64)),
                 f"train kernel {kernel}")
    # the standalone adjoint of the same batch
    g2 = _gpu_grads(cloud)
    d_rgba = torch.zeros((len(o), 4), dtype=torch.float32, device="cuda")
    d_rgba[:, :3] = 2.0 * (rgba[:, :3] - T)
    bv.adjoint(cloud, _dev(o), _dev(d), _dev(s), rgba, d_rgba, g2, settings=st, kernel=kernel)
    torch.cuda.synchronize()
    _check_grads(g2, _oracle_adjoint(scene, o, d, s, got_rgba, target.astype(np.float64)), f"adjoint {kernel}")


@pytest.mark.parametrize("n", [4096, 65536], ids=["direct", "packed"])
def test_c3_one_camera_shares_its_cells(shell, n):
    """One camera's rays at the 1M-cell scene: every ray starts in the
    camera's cell and most cross the same few cells next, whose gradients the
    fused kernel sums per block in LDS.  With the gradient tables added to
    directly (4,096 rays) and through the packed slots (65,536 rays, at least
    cells / 16), the fused training call and the fast adjoint equal the FP64
    oracle's adjoint.  Rays whose FP64 walk leaves a cell by another face than
    the FP32 walk (near-degenerate planes: 3 of the 4,096 here) are left out,
    since the kernels follow the FP32 walk bit for bit and the FP64 adjoint
    would sum other cells for them."""
    import torch

    import bv

    print(f"case C3 one camera, {n} rays", flush=True)
    scene, cloud, cams, starts = shell
    o, d, s = _camera_batch(cams, starts, [5], n, seed=41)
    osc = oracle.of_host(scene)
    w32 = oracle.trace_rays(osc, o, d, s, weight_threshold=0.001, max_steps=1024, depth=100.0)
    w64 = oracle.trace_rays(osc.as_dtype(np.float64), o, d, s, weight_threshold=0.001, max_steps=1024, depth=100.0,
                            dtype=np.float64)
    keep = (w32["steps"] == w64["steps"]) & (w32["last_point"] == w64["last_point"])
    assert keep.mean() > 0.99
    o, d, s = o[keep], d[keep], s[keep]
    n = len(o)
    target = np.random.default_rng(42).uniform(0, 1, size=(n, 3))
    st = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=100.0)
    T = _dev(target.astype(np.float32))
    grads = _gpu_grads(cloud)
    rgba = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    bv.train_step_l2(cloud, _dev(o), _dev(d), _dev(s), T, grads, settings=st, rgba_out=rgba)
    torch.cuda.synchronize()
    got_rgba = _host(rgba)
    ref = _oracle_adjoint(scene, o, d, s, got_rgba, target)
    _check_grads(grads, ref, f"one camera, fused, {n} rays")
    g2 = _gpu_grads(cloud)
    d_rgba = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    d_rgba[:, :3] = 2.0 * (rgba[:, :3] - T)
    bv.adjoint(cloud, _dev(o), _dev(d), _dev(s), rgba, d_rgba, g2, settings=st)
    torch.cuda.synchronize()
    _check_grads(g2, ref, f"one camera, adjoint, {n} rays")


def test_c4_gaussians_full_size():
    import torch

    import bv

    print("case C4: 524,288 Gaussians, 256 rays", flush=True)
    g = scenes.gaussian_scene(524_288, seed=3, sh_degree=3)
    cloud = bv.GaussianGpuCloud(g.mean, g.opacity, g.rotation, g.scale, g.sh, 3, 0.01)
    cam = scenes.looking_at((0.0, 0.4, 3.0), (0.0, 0.0, 0.0), math.radians(60.0), 1.0, 100.0)
    o, d = bv.camera_rays(cam, 1024, 1024)
    pix = np.random.default_rng(25).integers(0, 1024 * 1024, 256)
    o, d = _host(o)[pix], _host(d)[pix]
    st = bv.GaussianTraceSettings(0.01, 0.01, 0.0, 100.0)
    out = bv.gaussian_trace_rays(cloud, _dev(o), _dev(d), st, order_len=128)
    torch.cuda.synchronize()
    gsc = oracle.of_host(g)
    ref = oracle.trace_gaussians(gsc, o, d, min_opacity=0.01, min_transmittance=0.01, t_end=100.0, record=128)
    assert np.array_equal(_host(out["hits"], True), ref["hits"])
    assert ref["hits"].mean() > 5
    assert np.array_equal(_host(out["order"], True), ref["order"])
    err = float(np.abs(_host(out["rgba"]) - ref["rgba"]).max())
    assert err <= TOL, f"colour differs by {err}"
    d_rgba = np.random.default_rng(26).normal(size=(256, 4)).astype(np.float32)
    grads = bv.gaussian_zero_grads(cloud)
    bv.gaussian_adjoint(cloud, _dev(o), _dev(d), out["rgba"], _dev(d_rgba), grads, st)
    torch.cuda.synchronize()
    # At full size a few very thin particles make the FP32 adjoint formula
    # lose up to ~5e-4 of a table's largest entry against FP64 (mean
    # 1.5e-4, rotation 5.2e-4 for the oracle's own FP32 restatement), so the
    # device is held to the FP32 oracle at 1e-4 and to FP64 no worse than
    # the FP32 restatement itself.
    kw = dict(min_opacity=0.01, min_transmittance=0.01, t_end=100.0)
    ref32 = oracle.adjoint_gaussians(gsc, o, d, d_rgba, **kw)
    ref64 = oracle.adjoint_gaussians(gsc, o, d, d_rgba.astype(np.float64), dtype=np.float64, **kw)
    for k in ("mean", "scale", "rotation", "opacity", "sh"):
        gk = _host(grads[k]).reshape(ref64[k].shape).astype(np.float64)
        scale = float(np.abs(ref64[k]).max())
        assert float(np.abs(gk - ref32[k]).max()) <= 1e-4 * scale, k
        own = float(np.abs(ref32[k] - ref64[k]).max())
        assert float(np.abs(gk - ref64[k]).max()) <= own + 1e-4 * scale, k
    cloud.deinit()


def test_c5_two_million_site_splats():
    import torch

    import bv

    print("case C5: 2,097,152 PowerFoam sites, 256 rays", flush=True)
    h = scenes.powerfoam_scene(2_097_152, seed=4, sh_degree=0, cech=bv.compute_cech)
    cloud = bv.RadFoamGpuCloud(h)
    half = (2_097_152 / 65536) ** (1 / 3)
    cam = scenes.looking_at((0.0, 0.0, -2.4 * half), (0.0, 0.0, 0.0), math.radians(60.0), 16 / 9, 100.0)
    o, d = bv.camera_rays(cam, 1920, 1080)
    pix = np.random.default_rng(27).integers(0, 1920 * 1080, 256)
    o, d = _host(o)[pix], _host(d)[pix]
    L = 128
    rec = bv.splat_record(cloud, _dev(o), _dev(d), L, depth=100.0)
    torch.cuda.synchronize()
    osc = oracle.of_host(h)
    ref = oracle.record_splats(osc, o, d, max_steps=L, depth=100.0)
    cnt = _host(rec["count"], True)
    assert np.array_equal(cnt, ref["count"]) and cnt.mean() > 10
    live = np.arange(L)[None, :] < cnt[:, None].astype(np.int64)
    for k in ("cells", "nears", "dts"):
        a = _host(rec[k], k == "cells")
        assert np.array_equal(a[live], ref[k][live]), k
    st = bv.TraceSettings(weight_threshold=0.001, max_steps=1024, depth=100.0)
    tr = bv.splat_trace_rays(cloud, _dev(o), _dev(d), settings=st, outputs=("rgba", "steps"))
    torch.cuda.synchronize()
    tref = oracle.trace_splats(osc, o, d, 0, weight_threshold=0.001, max_steps=1024, depth=100.0)
    assert np.array_equal(_host(tr["steps"], True), tref["steps"])
    assert float(np.abs(_host(tr["rgba"]) - tref["rgba"]).max()) <= TOL
    jr = bv.splat_record_jacobians(cloud, _dev(o[:64]), _dev(d[:64]), L, depth=100.0)
    torch.cuda.synchronize()
    jref = oracle.record_splat_jacobians(osc, o[:64], d[:64], max_steps=L, depth=100.0)
    st_j = _host(jr["status"], True)
    assert np.array_equal(st_j, jref["status"])
    jl = np.arange(L)[None, :] < (st_j & 0x7FFFFFFF)[:, None].astype(np.int64)
    for k in ("cells", "next_cells", "previous_cells"):
        assert np.array_equal(_host(jr[k], True)[jl], jref[k][jl]), k
    assert np.array_equal(_host(jr["dts"])[jl], jref["dts"][jl])
    cloud.deinit()
