"""bv_volume_graph_step (the trainer graph's forward + backward) against the oracle.

Bar: loss, pixels, opacity and dt_from_positions within 1e-5 of the FP32
oracle (absolute; the loss relative 1e-5); every gradient table within 1e-4 of
its largest entry against the FP64 oracle (float atomics add in any order).
Cases: the reference's fixed-path fixture; recorded paths of a Delaunay scene
at SH 0 and 3 for every loss, ReLU and softplus, background, per-view exposure
and the opacity loss; the recorded and linearized dt of weighted paths;
status-trimmed rows equal full rows; all-padding rows add nothing; steps naming
cells outside the model are padding; and rows recorded by bv_record_paths on
the device, 4,096 of them, end to end.
"""
import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu

TOL_FWD = 1e-5
TOL_GRAD = 1e-4
GRAD_KEYS = ("d_log_density", "d_positions", "d_sh", "d_log_radii", "d_exposure")


def _dev(a, dtype=None):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    t = torch.from_numpy(a)
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _recorded(seed=4, n=400, deg=3, rays=48, L=96, depth=2.0):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    K = (deg + 1) ** 2
    host = scenes.HostScene(xyz, np.exp(rng.normal(size=n)).astype(np.float32),
                            rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32), deg, off, nb)
    o = rng.uniform(-0.3, 0.3, size=(rays, 3)).astype(np.float32)
    d = rng.normal(size=(rays, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    starts = np.array([scenes.containing_cell(host, p) for p in o], np.uint32)
    rec = oracle.record_paths(oracle.of_host(host), o, d, starts, max_steps=L, depth=depth, max_path_dt=50.0)
    cnt = (rec["status"] & 0x7FFFFFFF).astype(np.int64)
    mask = (np.arange(L)[None, :] < cnt[:, None]).astype(np.float32)
    b = dict(cells=np.where(mask > 0, rec["cells"], 0).astype(np.uint32),
             next_cells=np.where(mask > 0, rec["next_cells"], 0).astype(np.uint32),
             dts=np.where(mask > 0, rec["dts"], 0).astype(np.float32), mask=mask, ray_origin=o, ray_dir=d,
             labels=rng.uniform(0, 1, size=(rays, 3)).astype(np.float32), status=rec["status"])
    prm = dict(log_density=host.density.copy(), positions=host.xyz.copy(), sh=host.sh.copy(), sh_degree=deg)
    return host, prm, b


def _run_gpu(prm, b, opt, with_radii=False):
    import torch

    import bv

    n = prm["log_density"].shape[0]
    K = (prm["sh_degree"] + 1) ** 2
    dp = {k: (_dev(v) if isinstance(v, np.ndarray) else v) for k, v in prm.items()}
    db = {k: _dev(v.astype(np.float32) if np.asarray(v).dtype == np.float64 else v) for k, v in b.items()}
    z = lambda *s: torch.zeros(s, dtype=torch.float32, device="cuda")
    grads = dict(d_log_density=z(n), d_positions=z(n, 3), d_sh=z(n, 3 * K))
    if with_radii:
        grads["d_log_radii"] = z(n)
    if prm.get("exposure") is not None:
        grads["d_exposure"] = z(*prm["exposure"].shape)
    out = bv.volume_graph_step(dp, db, opt, grads)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res.update({k: v.cpu().numpy() for k, v in grads.items()})
    return res


def _compare(got, prm, b, kw, label):
    f32 = oracle.graph_step(prm, b, dtype=np.float32, **kw)
    f64 = oracle.graph_step({k: (np.asarray(v, np.float64) if isinstance(v, np.ndarray) and v.dtype == np.float32
                                 else v) for k, v in prm.items()},
                            {k: (np.asarray(v, np.float64) if isinstance(v, np.ndarray) and v.dtype == np.float32
                                 else v) for k, v in b.items()}, dtype=np.float64, **kw)
    assert abs(float(got["loss"][0]) - f32["loss"]) <= TOL_FWD * max(1.0, abs(f32["loss"])), label
    for k in ("pixels", "opacity", "dt"):
        err = float(np.abs(got[k] - f32[k]).max())
        assert err <= TOL_FWD, f"{label}: {k} differs by {err}"
    for k in GRAD_KEYS:
        if k not in got:
            continue
        ref = f64[k].reshape(got[k].shape)
        scale = float(np.abs(ref).max())
        err = float(np.abs(got[k] - ref).max())
        assert err <= TOL_GRAD * max(scale, 1e-12), f"{label}: {k} err {err} vs max {scale}"
    return f64


def test_reference_fixed_path_fixture():
    import bv

    pts = np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]], np.float32)
    prm = dict(log_density=np.full(4, 4.0, np.float32), positions=pts, sh=np.zeros((4, 3), np.float32),
               sh_degree=0)
    b = dict(cells=np.array([[0]], np.uint32), next_cells=np.array([[3]], np.uint32),
             dts=np.array([[1.25]], np.float32), mask=np.ones((1, 1), np.float32),
             ray_origin=np.array([[0.1, 0.1, -1.0]], np.float32), ray_dir=np.array([[0.0, 0.0, 1.0]], np.float32),
             labels=np.array([[0.1, 0.2, 0.3]], np.float32))
    got = _run_gpu(prm, b, bv.VolumeGraphOptions(loss=bv.LOSS_L1))
    _compare(got, prm, b, dict(loss=oracle.LOSS_L1), "fixture")
    assert got["dt"][0, 0] == pytest.approx(1.25, abs=1e-6)


@pytest.mark.parametrize("deg", [0, 3])
@pytest.mark.parametrize("loss", [0, 1, 2], ids=["l1", "smooth_l1", "l2"])
@pytest.mark.parametrize("beta", [0.0, 10.0], ids=["relu", "softplus"])
def test_positions_mode_matches_oracle(deg, loss, beta):
    import bv

    host, prm, b = _recorded(seed=10 + deg, deg=deg)
    rng = np.random.default_rng(3)
    if beta > 0:
        prm["log_density"] = (np.log(np.expm1(beta * prm["log_density"].astype(np.float64))) / beta).astype(np.float32)
    prm["exposure"] = rng.uniform(0.8, 1.2, size=(3, 3)).astype(np.float32)
    b["view_idx"] = rng.integers(0, 3, size=len(b["cells"])).astype(np.uint32)
    b["target_alpha"] = rng.uniform(0.5, 1.0, size=len(b["cells"])).astype(np.float32)
    kw = dict(loss=loss, softplus_beta=beta, background=(0.25, 0.0, 0.6), opacity_weight=0.2)
    got = _run_gpu(prm, b, bv.VolumeGraphOptions(**kw))
    _compare(got, prm, b, kw, f"deg{deg} loss{loss} beta{beta}")


@pytest.mark.parametrize("rays", [16, 480], ids=["direct", "packed"])
def test_positions_mode_direct_and_packed_tables(rays):
    """Density and position gradients go straight to the tables for batches
    of fewer rows than cells / 16 and through the packed per-cell slots for
    larger ones (400 cells: 16 rows direct, 480 rows packed); both equal the
    oracle."""
    import bv

    print(f"case graph positions {rays} rows", flush=True)
    host, prm, b = _recorded(seed=31, rays=rays)
    kw = dict(loss=oracle.LOSS_L2)
    got = _run_gpu(prm, b, bv.VolumeGraphOptions(**kw))
    _compare(got, prm, b, kw, f"{rays} rows")
    assert np.abs(got["d_positions"]).max() > 0 and np.abs(got["d_log_density"]).max() > 0


def test_basis_input_equals_device_basis():
    import bv

    host, prm, b = _recorded(seed=21, deg=3, rays=16)
    lib = oracle._load()
    import ctypes as C

    lib.bvo_graph_sh_basis_f32.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    basis = np.zeros((len(b["cells"]), 15), np.float32)
    for p in range(len(basis)):
        out = np.zeros(16, np.float32)
        lib.bvo_graph_sh_basis_f32(np.ascontiguousarray(b["ray_dir"][p]).ctypes.data, 16, out.ctypes.data)
        basis[p] = out[1:]
    opt = bv.VolumeGraphOptions(loss=bv.LOSS_SMOOTH_L1)
    a = _run_gpu(prm, b, opt)
    c = _run_gpu(prm, dict(b, basis=basis), opt)
    for k in ("pixels", "d_sh", "d_positions"):
        assert np.array_equal(a[k], c[k]) or float(np.abs(a[k] - c[k]).max()) <= 1e-7, k


@pytest.mark.parametrize("mode", [1, 2], ids=["recorded", "linearized"])
def test_weighted_modes_match_oracle(mode):
    import bv

    host, prm, b = _recorded(seed=30 + mode, deg=1)
    rng = np.random.default_rng(5)
    P, L = b["cells"].shape
    prm["log_radii"] = rng.uniform(-0.05, 0.05, size=host.n).astype(np.float32)
    b["previous_cells"] = np.concatenate([b["cells"][:, :1], b["cells"][:, :-1]], axis=1)
    for k in ("dt_grad_previous", "dt_grad_current", "dt_grad_next"):
        b[k] = (rng.normal(scale=0.3, size=(P, L, 4)) * b["mask"][..., None]).astype(np.float32)
    b["dt_reference_tangent"] = (rng.normal(scale=0.01, size=(P, L)) * b["mask"]).astype(np.float32)
    kw = dict(loss=oracle.LOSS_SMOOTH_L1, dt_mode=mode)
    got = _run_gpu(prm, b, bv.VolumeGraphOptions(**kw), with_radii=True)
    _compare(got, prm, b, kw, f"mode{mode}")
    if mode == 1:
        assert not got["d_positions"].any() and not got["d_log_radii"].any()
    else:
        assert np.abs(got["d_log_radii"]).max() > 0


def test_status_rows_padding_and_foreign_cells():
    import bv

    host, prm, b = _recorded(seed=40, deg=3)
    opt = bv.VolumeGraphOptions(loss=bv.LOSS_L1)
    full = dict(b)
    full.pop("status")
    a = _run_gpu(prm, full, opt)
    c = _run_gpu(prm, b, opt)
    for k in ("pixels", "dt"):
        assert float(np.abs(a[k] - c[k]).max()) <= 1e-6, k
    for k in ("d_log_density", "d_positions", "d_sh"):
        assert float(np.abs(a[k] - c[k]).max()) <= 1e-6 * max(1.0, float(np.abs(a[k]).max())), k
    # all padding: zero loss against zero labels, nothing added
    pad = dict(full, mask=np.zeros_like(b["mask"]), labels=np.zeros_like(b["labels"]))
    z = _run_gpu(prm, pad, opt)
    assert float(z["loss"][0]) == 0.0
    for k in ("d_log_density", "d_positions", "d_sh"):
        assert not z[k].any(), k
    # steps naming cells beyond the model are padding (rows after them keep their own values)
    bad = dict(full)
    bad["cells"] = b["cells"].copy()
    bad["mask"] = b["mask"].copy()
    bad["cells"][:, 60:] = host.n + 7
    bad["mask"][:, 60:] = 1.0
    ref = dict(full)
    ref["mask"] = b["mask"].copy()
    ref["mask"][:, 60:] = 0.0
    x = _run_gpu(prm, bad, opt)
    y = _run_gpu(prm, ref, opt)
    assert float(np.abs(x["pixels"] - y["pixels"]).max()) <= 1e-6


def test_device_recorded_rows_end_to_end():
    """4,096 rays: bv_record_paths on the device (rows bit-exact to the oracle's
    recorder), then the graph on those rows."""
    import torch

    import bv

    rng = np.random.default_rng(50)
    n, deg, P, L = 2000, 3, 4096, 128
    xyz = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    K = (deg + 1) ** 2
    host = scenes.HostScene(xyz, np.exp(rng.normal(size=n)).astype(np.float32),
                            rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32), deg, off, nb)
    cloud = bv.RadFoamGpuCloud(host)
    o = rng.uniform(-0.5, 0.5, size=(P, 3)).astype(np.float32)
    d = rng.normal(size=(P, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    starts = np.array([scenes.containing_cell(host, p) for p in o], np.uint32)
    rec = bv.PathRecorder(cloud)
    bufs = rec.buffers(P, L)
    rec.dispatch(_dev(o), _dev(d), _dev(starts.view(np.int32)), max_steps=L, depth=3.0, buffers=bufs)
    torch.cuda.synchronize()
    status = bufs.status.cpu().numpy().view(np.uint32)
    ref_rec = oracle.record_paths(oracle.of_host(host), o, d, starts, max_steps=L, depth=3.0, max_path_dt=50.0)
    assert np.array_equal(status, ref_rec["status"])
    cnt = (status & 0x7FFFFFFF).astype(np.int64)
    live = np.arange(L)[None, :] < cnt[:, None]
    cells = bufs.cells.cpu().numpy().view(np.uint32)
    assert np.array_equal(cells[live], ref_rec["cells"][live])
    mask = bufs.mask.cpu().numpy()
    assert np.all(mask[~live] == 0.0) and np.all(mask[live] == 1.0)
    labels = rng.uniform(0, 1, size=(P, 3)).astype(np.float32)
    b = dict(cells=cells, next_cells=bufs.next_cells.cpu().numpy().view(np.uint32), dts=bufs.dts.cpu().numpy(),
             mask=mask, status=status, ray_origin=o, ray_dir=d, labels=labels)
    prm = dict(log_density=host.density, positions=host.xyz, sh=host.sh, sh_degree=deg)
    got = _run_gpu(prm, b, bv.VolumeGraphOptions(loss=bv.LOSS_SMOOTH_L1))
    _compare(got, prm, b, dict(loss=oracle.LOSS_SMOOTH_L1), "device rows")
    cloud.deinit()
