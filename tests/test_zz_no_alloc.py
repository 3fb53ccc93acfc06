"""bv_train_step_l2_noalloc against bv_train_step_l2 (BV_KERNEL_FAST).

The allocation-free entry runs the same fused kernel over the library's fixed
per-GPU workspace, without the packed gradient table, so colours must agree
bit for bit and the loss and gradients up to FP32 summation order (relative
1e-5 of each table's largest entry).  Two calls in a row, the second smaller,
and once more with an 8-step path cache so that the re-walk replay runs too.
No device memory may be allocated by the call (free HBM unchanged).
"""
import ctypes as C

import numpy as np
import pytest

from bv import scenes

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.4, 0.4, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = np.array([scenes.containing_cell(scene, p) for p in o], np.uint32)
    target = rng.uniform(0, 1, size=(n, 3)).astype(np.float32)
    return _dev(o), _dev(d), _dev(starts.view(np.int32)), _dev(target)


def _close(a, b, label):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    scale = max(float(np.abs(b).max()), 1e-30)
    err = float(np.abs(a - b).max())
    assert err <= 1e-5 * scale + 1e-7, f"{label}: {err} vs max {scale}"


@pytest.mark.parametrize("cache_steps", [None, "8"], ids=["cache", "rewalk"])
def test_noalloc_training_matches_the_allocating_entry(small_scene, cache_steps, monkeypatch):
    import torch

    import bv

    if cache_steps is not None:
        monkeypatch.setenv("BV_PATH_CACHE_STEPS", cache_steps)
    cloud = bv.RadFoamGpuCloud(small_scene)
    st = bv.TraceSettings(weight_threshold=1e-3, max_steps=1024, depth=10.0)
    for n, seed in ((3000, 1), (700, 2)):
        o, d, s, target = _rays(small_scene, n, seed)
        ga, gb = bv.zero_grads(cloud), bv.zero_grads(cloud)
        ra = torch.empty((n, 4), dtype=torch.float32, device="cuda")
        rb = torch.empty_like(ra)
        la = torch.zeros(1, dtype=torch.float32, device="cuda")
        lb = torch.zeros_like(la)
        bv.train_step_l2(cloud, o, d, s, target, ga, settings=st, rgba_out=ra, loss=la, kernel=bv.KERNEL_FAST)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        bv.train_step_l2_noalloc(cloud, o, d, s, target, gb, settings=st, rgba_out=rb, loss=lb)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0, "the allocation-free entry changed free device memory"
        assert torch.equal(ra, rb), f"{n} rays: colours differ"
        _close(lb, la, f"{n} rays: loss")
        for k in ("density", "sh", "xyz"):
            assert float(ga[k].abs().max()) > 0.0
            _close(gb[k], ga[k], f"{n} rays: d_{k}")
    cloud.deinit()


def test_noalloc_rejects_what_it_cannot_take(small_scene):
    import bv
    from bv import _lib

    cloud = bv.RadFoamGpuCloud(small_scene)
    lib = _lib.load()
    st = bv.TraceSettings()
    rays = _lib.Rays(1, 1, None, 0, 0, 1 << 31)
    g = _lib.Grads(1, 1, 1)
    rc = lib.bv_train_step_l2_noalloc(cloud.handle, C.byref(rays), C.byref(st.c()), 1, None, C.byref(g), None,
                                      None)
    assert rc == _lib.BV_ERR_INVALID and b"2^31" in lib.bv_last_error()
    cloud.deinit()
