"""bv_splat_record_jacobians (the PowerFoam recorder's Jacobian streams) against
the oracle, and the weighted trainer graph on the streams it records.

Bar: cells, next / previous cells, dt and status bit-exact; the three
Jacobian rows and the reference tangent within 1e-5 relative (floored at 1)
of the oracle's FP32 values.  Then bv_volume_graph_step in the linearized dt
mode on the device-recorded rows matches the oracle graph on the same rows,
and at the recorded geometry its dt_from_positions reproduces the recorded dt
(the tangent cancels) within 1e-5.
"""
import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu

JAC = ("dt_grad_previous", "dt_grad_current", "dt_grad_next")


def _dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def _host(t, u32=False):
    a = t.cpu().numpy()
    return a.view(np.uint32) if u32 else a


def _scene(n, seed, sh_degree=1):
    return scenes.powerfoam_scene(n, seed=seed, sh_degree=sh_degree)


def _rays(n, rng, half=1.0):
    o = np.column_stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n),
                         np.full(n, -2.6 * half)]).astype(np.float32)
    d = np.column_stack([rng.uniform(-0.35, 0.35, n), rng.uniform(-0.35, 0.35, n), np.ones(n)]).astype(np.float32)
    return o, d


def _compare(got, ref, L, label):
    st = _host(got["status"], True)
    np.testing.assert_array_equal(st, ref["status"], err_msg=f"{label}: status")
    cnt = st & 0x7FFFFFFF
    live = np.arange(L)[None, :] < cnt[:, None]
    for k in ("cells", "next_cells", "previous_cells"):
        np.testing.assert_array_equal(_host(got[k], True)[live], ref[k][live], err_msg=f"{label}: {k}")
    np.testing.assert_array_equal(_host(got["dts"])[live], ref["dts"][live], err_msg=f"{label}: dts")
    assert np.all(_host(got["mask"])[live] == 1.0)
    for k in JAC + ("dt_reference_tangent",):
        a, b = _host(got[k])[live], ref[k][live]
        err = np.abs(a - b) / np.maximum(1.0, np.abs(b))
        assert float(err.max()) <= 1e-5, f"{label}: {k} rel err {float(err.max())}"
    return live


@pytest.mark.parametrize("L,max_dt", [(128, 50.0), (6, 50.0), (128, 0.01)], ids=["full", "truncated", "clamped"])
def test_jacobian_rows_match_oracle(L, max_dt):
    import bv

    h = _scene(4096, seed=7)
    cloud = bv.RadFoamGpuCloud(h)
    rng = np.random.default_rng(11)
    o, d = _rays(512, rng, half=(4096 / 65536) ** (1 / 3))
    got = bv.splat_record_jacobians(cloud, _dev(o), _dev(d), L, depth=10.0, max_path_dt=max_dt)
    ref = oracle.record_splat_jacobians(oracle.of_host(h), o, d, max_steps=L, depth=10.0, max_path_dt=max_dt)
    live = _compare(got, ref, L, f"L{L} max_dt{max_dt}")
    assert live.sum() > 1000
    if L == 6:
        assert np.any(ref["status"] >> 31)
    if max_dt < 1:
        assert np.any(ref["dts"][live] == np.float32(max_dt))
    # same cells and dt as the recorder without Jacobians (when nothing is clamped)
    if max_dt > 1:
        plain = bv.splat_record(cloud, _dev(o), _dev(d), L, depth=10.0)
        assert np.array_equal(_host(plain["cells"], True)[live], _host(got["cells"], True)[live])
        assert np.array_equal(_host(plain["dts"])[live], _host(got["dts"])[live])
    cloud.deinit()


def test_weighted_graph_on_device_recorded_streams():
    import torch

    import bv

    h = _scene(4096, seed=9, sh_degree=1)
    cloud = bv.RadFoamGpuCloud(h)
    rng = np.random.default_rng(12)
    P, L = 1024, 96
    o, d = _rays(P, rng, half=(4096 / 65536) ** (1 / 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rec = bv.splat_record_jacobians(cloud, _dev(o), _dev(d), L, depth=10.0)
    torch.cuda.synchronize()
    b = {k: (_host(v, True) if v.dtype == torch.int32 else _host(v)) for k, v in rec.items()}
    b.update(ray_origin=o, ray_dir=d, labels=rng.uniform(0, 1, size=(P, 3)).astype(np.float32))
    beta = 100.0
    # log radius whose softplus_100 is the recorded radius
    log_r = (np.log(np.expm1(beta * h.radius.astype(np.float64))) / beta).astype(np.float32)
    prm = dict(log_density=h.density, positions=h.xyz, sh=h.sh, sh_degree=1, log_radii=log_r)
    opt = bv.VolumeGraphOptions(loss=bv.LOSS_SMOOTH_L1, dt_mode=bv.GRAPH_DT_LINEARIZED)
    dp = {k: (_dev(v) if isinstance(v, np.ndarray) else v) for k, v in prm.items()}
    n = h.n
    z = lambda *s: torch.zeros(s, dtype=torch.float32, device="cuda")
    grads = dict(d_log_density=z(n), d_positions=z(n, 3), d_sh=z(n, 3 * 4), d_log_radii=z(n))
    out = bv.volume_graph_step(dp, {k: v for k, v in rec.items()} | dict(ray_origin=_dev(o), ray_dir=_dev(d),
                                                                        labels=_dev(b["labels"])), opt, grads)
    torch.cuda.synchronize()
    live = (np.arange(L)[None, :] < (b["status"] & 0x7FFFFFFF)[:, None])
    assert live.sum() > 2000
    # at the recorded geometry the linearization returns the recorded dt
    err = np.abs(out["dt"].cpu().numpy()[live] - b["dts"][live]) / np.maximum(1.0, np.abs(b["dts"][live]))
    assert float(err.max()) <= 1e-5
    ref = oracle.graph_step({k: (v.astype(np.float64) if isinstance(v, np.ndarray) else v) for k, v in prm.items()},
                            {k: (v.astype(np.float64) if isinstance(v, np.ndarray) and v.dtype == np.float32 else v)
                             for k, v in b.items()}, loss=oracle.LOSS_SMOOTH_L1, dt_mode=oracle.DT_LINEARIZED,
                            dtype=np.float64)
    for k in ("d_log_density", "d_positions", "d_sh", "d_log_radii"):
        g = grads[k].cpu().numpy().reshape(ref[k].shape)
        scale = float(np.abs(ref[k]).max())
        assert scale > 0 and float(np.abs(g - ref[k]).max()) <= 1e-4 * scale, k
    assert abs(float(out["loss"][0]) - ref["loss"]) <= 1e-5 * max(1.0, ref["loss"])
    cloud.deinit()
