"""Caller mistakes the library must absorb without reading outside its buffers.

* A per-ray start cell outside the scene (BV_NO_CELL or n + 5) makes that ray
  empty in every walk entry point: zero colour, zero steps, status 0, no
  gradient, while the other rays of the batch are unchanged.
* Containing-cell queries of one scene issued from several host threads, each
  on its own stream, each get their own answer (the brute force of
  lib.rs:526-541).
* A call of 2^32 or more rays is rejected before any launch.
"""
import threading

import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _batch(scene, n=64, seed=3):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.3, 0.3, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    starts = np.array([scenes.containing_cell(scene, p) for p in o], np.uint32)
    bad = np.array([1, 7, 20, 41, 63])
    starts[bad[::2]] = 0xFFFFFFFF
    starts[bad[1::2]] = scene.n + 5
    good = np.setdiff1d(np.arange(n), bad)
    return o, d, starts, bad, good


@pytest.mark.parametrize("kernel", [0, 1], ids=["plain", "fast"])
def test_out_of_range_start_cells_give_empty_rays(small_scene, kernel):
    import bv

    cloud = bv.RadFoamGpuCloud(small_scene)
    o, d, starts, bad, good = _batch(small_scene)
    st = bv.TraceSettings(weight_threshold=1e-3, max_steps=512, depth=10.0)
    got = bv.trace_rays(cloud, _dev(o), _dev(d), _dev(starts.view(np.int32)), settings=st, kernel=kernel,
                        outputs=("rgba", "steps", "t_end"))
    rgba = got["rgba"].cpu().numpy()
    assert np.all(rgba[bad] == 0.0)
    assert np.all(got["steps"].cpu().numpy()[bad] == 0)
    assert np.all(got["t_end"].cpu().numpy()[bad] == 0.0)
    ref = oracle.trace_rays(oracle.of_host(small_scene), o[good], d[good], starts[good], weight_threshold=1e-3,
                            max_steps=512, depth=10.0)
    assert np.array_equal(got["steps"].cpu().numpy().view(np.uint32)[good], ref["steps"])
    assert float(np.abs(rgba[good] - ref["rgba"]).max()) <= 1e-5
    cloud.deinit()


def test_out_of_range_start_cells_record_nothing_and_add_no_gradient(small_scene):
    import torch

    import bv

    cloud = bv.RadFoamGpuCloud(small_scene)
    o, d, starts, bad, good = _batch(small_scene)
    rec = bv.PathRecorder(cloud)
    bufs = rec.dispatch(_dev(o), _dev(d), _dev(starts.view(np.int32)), max_steps=64, depth=10.0)
    status = bufs.status.cpu().numpy().view(np.uint32)
    assert np.all(status[bad] == 0)
    ref = oracle.record_paths(oracle.of_host(small_scene), o[good], d[good], starts[good], max_steps=64, depth=10.0)
    assert np.array_equal(status[good], ref["status"])
    # only the bad rays: every gradient stays exactly zero, in the standalone
    # adjoint and in the fused training kernel (whose empty rays still add
    # their squared target to the loss)
    st = bv.TraceSettings(weight_threshold=1e-3, max_steps=512, depth=10.0)
    ob, db, sb = _dev(o[bad]), _dev(d[bad]), _dev(starts[bad].view(np.int32))
    for kernel in (0, 1):
        grads = bv.zero_grads(cloud)
        rgba = torch.zeros((len(bad), 4), dtype=torch.float32, device="cuda")
        g = torch.ones((len(bad), 4), dtype=torch.float32, device="cuda")
        bv.adjoint(cloud, ob, db, sb, rgba, g, grads, settings=st, kernel=kernel)
        target = torch.full((len(bad), 3), 0.5, dtype=torch.float32, device="cuda")
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        out = torch.full((len(bad), 4), 7.0, dtype=torch.float32, device="cuda")
        bv.train_step_l2(cloud, ob, db, sb, target, grads, settings=st, rgba_out=out, loss=loss, kernel=kernel)
        torch.cuda.synchronize()
        for k, v in grads.items():
            assert float(v.abs().max()) == 0.0, (kernel, k)
        assert np.all(out.cpu().numpy() == 0.0)
        assert abs(float(loss.item()) - 0.75 * len(bad)) <= 1e-5
    cloud.deinit()


def test_containing_cell_queries_on_several_streams(small_scene):
    import torch

    import bv

    cloud = bv.RadFoamGpuCloud(small_scene)
    rng = np.random.default_rng(8)
    queries = rng.uniform(-1.2, 1.2, size=(4, 24, 3)).astype(np.float32)
    want = [[scenes.containing_cell(small_scene, q) for q in qs] for qs in queries]
    got = [[None] * len(qs) for qs in queries]
    errors = []

    def worker(t):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for k, q in enumerate(queries[t]):
                    got[t][k] = cloud.containing_point(q)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(len(queries))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    assert got == want
    cloud.deinit()


def test_ray_count_beyond_32_bits_is_rejected(small_scene):
    """A call of 2^32 or more rays fails validation (BV_ERR_INVALID) before any
    launch: the persistent kernels count claimed rays in 32 bits."""
    import ctypes as C

    import torch

    import bv

    print("case guard ray count 2^32", flush=True)
    cloud = bv.RadFoamGpuCloud(small_scene)
    o = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    d = torch.ones((4, 3), dtype=torch.float32, device="cuda")
    rays = bv._rays(o, d, None, 0)
    rays.count = 1 << 32
    out = bv._alloc_outputs(4, ("rgba",), o.device)
    st = bv.TraceSettings().c()
    rc = cloud._lib.bv_trace_rays(cloud.handle, C.byref(rays), C.byref(st), C.byref(bv._outputs_c(out)), 1, None)
    assert rc == 1
    grads = bv.zero_grads(cloud)
    g = bv._lib.Grads(grads["density"].data_ptr(), grads["sh"].data_ptr(), grads["xyz"].data_ptr())
    target = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    rc = cloud._lib.bv_train_step_l2(cloud.handle, C.byref(rays), C.byref(st), target.data_ptr(), None,
                                     C.byref(g), None, 1, None)
    assert rc == 1
    rays.count = 4  # the same call with its real count runs
    rc = cloud._lib.bv_trace_rays(cloud.handle, C.byref(rays), C.byref(st), C.byref(bv._outputs_c(out)), 1, None)
    torch.cuda.synchronize()
    assert rc == 0
    cloud.deinit()
