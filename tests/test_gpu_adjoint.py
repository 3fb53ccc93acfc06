"""HIP adjoint against the oracle's adjoint and against FP64 central finite
differences (relative tolerance 1e-4 on sampled parameters, as stated for the
training workload), for both the plain and the fast kernel."""
import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[(0, None), (1, 1), (1, 4), (1, 16), (1, 32)],
                ids=["plain", "fast-g1", "fast-g4", "fast-g16", "fast-g32"])
def kernel(request, monkeypatch):
    k, lanes = request.param
    if lanes is not None:
        monkeypatch.setenv("BV_LANES_PER_RAY", str(lanes))
    return k


def _setup(scene, n, seed, depth):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.4, 0.4, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = np.array([scenes.containing_cell(scene, x) for x in o], np.uint32)
    target = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    return o, d, starts, target


def _gpu_adjoint(scene, o, d, starts, target, kernel, depth, wt):
    import torch

    import bv

    cloud = bv.RadFoamGpuCloud(scene)
    O, D = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    S = torch.from_numpy(starts.view(np.int32)).cuda()
    st = bv.TraceSettings(weight_threshold=wt, max_steps=1024, depth=depth)
    fwd = bv.trace_rays(cloud, O, D, S, settings=st, kernel=kernel)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    g = bv.l2_loss_grad(fwd["rgba"], torch.from_numpy(target).cuda(), loss=loss)
    grads = bv.zero_grads(cloud)
    bv.adjoint(cloud, O, D, S, fwd["rgba"], g, grads, settings=st, kernel=kernel)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in grads.items()}
    out["rgba"] = fwd["rgba"].cpu().numpy()
    out["d_rgba"] = g.cpu().numpy()
    out["loss"] = float(loss.item())
    cloud.deinit()
    return out


def _fd_loss(s64, o, d, starts, target, depth, wt):
    r = oracle.trace_rays(s64, o, d, starts, weight_threshold=wt, max_steps=1024, depth=depth, dtype=np.float64)
    e = r["rgba"][:, :3] - target
    return float((e * e).sum())


def test_adjoint_matches_oracle_and_fd(small_scene, kernel):
    print(f"case adjoint small kernel={kernel}", flush=True)
    depth, wt = 3.0, 0.001
    o, d, s, target = _setup(small_scene, 256, 1, depth)
    got = _gpu_adjoint(small_scene, o, d, s, target, kernel, depth, wt)
    ref_fwd = oracle.trace_rays(oracle.of_host(small_scene), o, d, s, weight_threshold=wt, depth=depth)
    assert np.abs(got["rgba"] - ref_fwd["rgba"]).max() <= 1e-5
    # loss gradient helper: 2 (rgb - target), 0 for alpha; loss = sum of squares
    np.testing.assert_allclose(got["d_rgba"][:, :3], 2 * (got["rgba"][:, :3] - target), rtol=1e-6, atol=1e-7)
    assert np.all(got["d_rgba"][:, 3] == 0)
    assert abs(got["loss"] - float(((got["rgba"][:, :3] - target) ** 2).sum())) <= 1e-4 * got["loss"]
    s64 = oracle.of_host(small_scene).as_dtype(np.float64)
    g64 = np.concatenate([got["d_rgba"][:, :3].astype(np.float64), np.zeros((len(o), 1))], 1)
    ref = oracle.adjoint(s64, o, d, s, g64, weight_threshold=wt, depth=depth, dtype=np.float64)
    for k in ("density", "sh", "xyz"):
        scale = np.abs(ref[k]).max()
        err = np.abs(got[k].reshape(ref[k].shape) - ref[k]).max()
        assert err <= 1e-4 * scale, (k, err, scale)
    # 64 sampled parameters against FP64 central differences of the FP64 forward.
    rng = np.random.default_rng(5)
    touched = np.flatnonzero(ref["density"] != 0)
    h = 1e-6
    checked = 0
    for cell in rng.choice(touched, size=22, replace=False):
        for kind, idx in (("density", (cell,)), ("sh", (cell, int(rng.integers(0, 48)))),
                          ("xyz", (cell, int(rng.integers(0, 3))))):
            arr = getattr(s64, kind)
            base = arr[idx]
            arr[idx] = base + h
            lp = _fd_loss(s64, o, d, s, target, depth, wt)
            arr[idx] = base - h
            lm = _fd_loss(s64, o, d, s, target, depth, wt)
            arr[idx] = base
            fd = (lp - lm) / (2 * h)
            an = float(got[kind].reshape(ref[kind].shape)[idx])
            # relative 1e-4, with a floor at 1e-4 of this kind's largest gradient
            floor = 1e-4 * np.abs(ref[kind]).max()
            assert abs(an - fd) <= max(1e-4 * abs(fd), floor), (kind, idx, an, fd)
            checked += 1
    assert checked >= 64


def test_adjoint_c1_batch(c1_scene, kernel):
    """Config-1 scene, a 4,096-ray batch of random rays: GPU vs oracle FP64 adjoint."""
    print(f"case adjoint c1 kernel={kernel}", flush=True)
    depth, wt = 4.0, 0.001
    sc = scenes.HostScene(c1_scene.xyz, c1_scene.density, c1_scene.sh, 0, c1_scene.offsets, c1_scene.neighbors)
    o, d, s, target = _setup(sc, 4096, 2, depth)
    got = _gpu_adjoint(sc, o, d, s, target, kernel, depth, wt)
    g64 = np.concatenate([got["d_rgba"][:, :3].astype(np.float64), np.zeros((len(o), 1))], 1)
    ref = oracle.adjoint(oracle.of_host(sc).as_dtype(np.float64), o, d, s, g64, weight_threshold=wt, depth=depth,
                         dtype=np.float64)
    for k in ("density", "sh", "xyz"):
        scale = np.abs(ref[k]).max()
        err = np.abs(got[k].reshape(ref[k].shape) - ref[k]).max()
        assert err <= 1e-4 * scale, (k, err, scale)


def test_autograd_op_matches_adjoint(small_scene):
    """RadFoamRender (the trainer's custom op) backward equals bv_adjoint."""
    import torch

    import bv

    depth, wt = 3.0, 0.001
    o, d, s, target = _setup(small_scene, 128, 3, depth)
    cloud = bv.RadFoamGpuCloud(small_scene)
    dens = torch.from_numpy(small_scene.density.copy()).cuda().requires_grad_(True)
    sh = torch.from_numpy(small_scene.sh.copy()).cuda().requires_grad_(True)
    xyz = torch.from_numpy(small_scene.xyz.copy()).cuda().requires_grad_(True)
    O, D = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    S = torch.from_numpy(s.view(np.int32)).cuda()
    st = bv.TraceSettings(weight_threshold=wt, depth=depth)
    rgba = bv.RadFoamRender.apply(cloud, dens, sh, xyz, O, D, S, st)
    loss = ((rgba[:, :3] - torch.from_numpy(target).cuda()) ** 2).sum()
    loss.backward()
    ref = _gpu_adjoint(small_scene, o, d, s, target, 1, depth, wt)
    for name, t in (("density", dens), ("sh", sh), ("xyz", xyz)):
        r = ref[name].reshape(t.shape)
        err = np.abs(t.grad.cpu().numpy() - r).max()
        assert err <= 1e-5 * max(np.abs(r).max(), 1.0), (name, err)
    cloud.deinit()


def test_train_step_l2_matches_separate_calls(small_scene, kernel):
    """bv_train_step_l2 (fused for the fast kernel) equals trace + L2 gradient +
    adjoint, and the FP64 oracle's adjoint of the same loss."""
    print(f"case train_step kernel={kernel}", flush=True)
    _check_train_step(small_scene, kernel)


@pytest.mark.parametrize("cache_steps", ["8", "0"], ids=["cache8-mixed", "cache-off"])
@pytest.mark.parametrize("lanes", ["16", "4"])
def test_train_step_l2_walks_beyond_the_path_cache(small_scene, cache_steps, lanes, monkeypatch):
    """The fused kernel replays a walk held in its path cache step-parallel and
    walks a longer one again.  With an 8-step cache a batch (and a wave) mixes
    both replays; with the cache off every walk is walked again.  Either way
    the gradients equal the separate trace + adjoint calls and the FP64
    oracle's adjoint."""
    import torch

    import bv

    print(f"case train_step cache={cache_steps} lanes={lanes}", flush=True)
    monkeypatch.setenv("BV_PATH_CACHE_STEPS", cache_steps)
    monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    o, d, s, _ = _setup(small_scene, 1000, 7, 3.0)
    cloud = bv.RadFoamGpuCloud(small_scene)
    st = bv.TraceSettings(weight_threshold=0.001, depth=3.0)
    steps = bv.trace_rays(cloud, torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda(),
                          torch.from_numpy(s.view(np.int32)).cuda(), settings=st, outputs=("rgba", "steps"))["steps"]
    steps = steps.cpu().numpy().view(np.uint32)
    cloud.deinit()
    if cache_steps == "8":
        assert (steps > 8).mean() > 0.2 and (steps <= 8).mean() > 0.2  # both replays occur
    _check_train_step(small_scene, 1)


@pytest.mark.parametrize("lanes", ["16", "4"])
@pytest.mark.parametrize("cache_steps", ["256", "8"])
def test_train_step_l2_one_launch_carries_many_batches(small_scene, lanes, cache_steps, monkeypatch):
    """One call with more rays than the device holds groups for (64,000 rays
    against 16,384 resident groups at 16 lanes per ray) runs as one
    persistent launch whose waves claim rays until they run out, each group
    reusing its path-cache region (and, with an 8-step cache, mixing cached
    and walked-again replays).  It equals 16 calls of 4,000 rays: colour
    bit-exact, loss and gradients to the rounding of the atomic sums."""
    import torch

    import bv

    print(f"case train_step one launch lanes={lanes} cache={cache_steps}", flush=True)
    monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    monkeypatch.setenv("BV_PATH_CACHE_STEPS", cache_steps)
    o0, _, s0, _ = _setup(small_scene, 1000, 9, 3.0)
    reps = 64
    rng = np.random.default_rng(3)
    o = np.tile(o0, (reps, 1))
    s = np.tile(s0, reps)
    d = rng.normal(size=(len(o), 3)).astype(np.float32)
    target = rng.uniform(0, 1, size=(len(o), 3)).astype(np.float32)
    cloud = bv.RadFoamGpuCloud(small_scene)
    O, D = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    S = torch.from_numpy(s.view(np.int32)).cuda()
    T = torch.from_numpy(target).cuda()
    st = bv.TraceSettings(weight_threshold=0.001, depth=3.0)
    res = []
    for chunk in (len(o), 4000):
        grads = bv.zero_grads(cloud)
        rgba = torch.empty((len(o), 4), dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        for a in range(0, len(o), chunk):
            b = a + chunk
            bv.train_step_l2(cloud, O[a:b], D[a:b], S[a:b], T[a:b], grads, settings=st, rgba_out=rgba[a:b],
                             loss=loss, kernel=1)
        torch.cuda.synchronize()
        res.append(({k: v.cpu().numpy() for k, v in grads.items()}, rgba.cpu().numpy(), float(loss.item())))
    cloud.deinit()
    (g1, c1, l1), (g2, c2, l2) = res
    np.testing.assert_array_equal(c1, c2)
    assert abs(l1 - l2) <= 1e-5 * l2
    for k in g1:
        scale = np.abs(g2[k]).max()
        assert scale > 0 and np.abs(g1[k] - g2[k]).max() <= 1e-5 * scale, k


def test_train_step_l2_small_call_adds_directly(small_scene):
    """A call with few rays for its scene (24 rays, 512 cells) adds density and
    position gradients straight into the caller's tables; larger calls go
    through the packed per-cell table.  Both equal the separate calls and the
    FP64 oracle."""
    print("case train_step 24 rays", flush=True)
    _check_train_step(small_scene, 1, n=24)


@pytest.mark.parametrize("n", [2048, 24], ids=["packed", "direct"])
@pytest.mark.parametrize("lanes", ["16", "4"])
def test_train_step_l2_rays_sharing_a_start_cell(small_scene, n, lanes, monkeypatch):
    """A fan of rays from one point, as a camera's rays are: every ray's first
    steps cross the same few cells, whose gradients the fused kernel sums per
    block in LDS before adding them to the tables (with the packed per-cell
    table and without).  Equal to the separate calls and the FP64 oracle."""
    print(f"case train_step shared start n={n} lanes={lanes}", flush=True)
    monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    rng = np.random.default_rng(21)
    o = np.tile(np.array([[0.05, -0.1, 0.02]], np.float32), (n, 1))
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:, 2] = np.abs(d[:, 2]) + 1.0  # a cone of directions, as a camera's
    s = np.full(n, scenes.containing_cell(small_scene, o[0]), np.uint32)
    target = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    _check_train_step(small_scene, 1, rays=(o, d, s, target))


def _check_train_step(small_scene, kernel, n=1000, rays=None):
    import torch

    import bv

    depth, wt = 3.0, 0.001
    o, d, s, target = rays if rays is not None else _setup(small_scene, n, 7, depth)
    cloud = bv.RadFoamGpuCloud(small_scene)
    O, D = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    S = torch.from_numpy(s.view(np.int32)).cuda()
    T = torch.from_numpy(target).cuda()
    st = bv.TraceSettings(weight_threshold=wt, depth=depth)
    grads = bv.zero_grads(cloud)
    rgba = torch.empty((len(o), 4), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    bv.train_step_l2(cloud, O, D, S, T, grads, settings=st, rgba_out=rgba, loss=loss, kernel=kernel)
    # a second batch accumulates: gradients double
    bv.train_step_l2(cloud, O, D, S, T, grads, settings=st, kernel=kernel)
    torch.cuda.synchronize()
    sep = _gpu_adjoint(small_scene, o, d, s, target, kernel, depth, wt)
    assert np.abs(rgba.cpu().numpy() - sep["rgba"]).max() <= 1e-6
    assert abs(loss.item() - sep["loss"]) <= 1e-5 * sep["loss"]
    g64 = np.concatenate([sep["d_rgba"][:, :3].astype(np.float64), np.zeros((len(o), 1))], 1)
    ref = oracle.adjoint(oracle.of_host(small_scene).as_dtype(np.float64), o, d, s, g64, weight_threshold=wt,
                         depth=depth, dtype=np.float64)
    for k in ("density", "sh", "xyz"):
        got = grads[k].cpu().numpy().reshape(ref[k].shape) / 2
        scale = np.abs(ref[k]).max()
        assert np.abs(got - ref[k]).max() <= 1e-4 * scale, k
        assert np.abs(got - sep[k].reshape(ref[k].shape)).max() <= 1e-5 * scale, k
    cloud.deinit()
