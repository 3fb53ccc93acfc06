"""The oracle of the trainer's volumetric graph (oracle/bv_graph_oracle.cpp).

Pins, each chosen so a dropped term, a wrong sign or a swapped operand fails:
  * the reference's fixed-path fixture (diff_render.rs:11961-12073): cell 0 ->
    cell 3 of the tiny model crosses the bisector z = 0.25 at t = 1.25, so the
    loss and every position gradient have closed forms;
  * the reference's weighted-linearization fixture (:12076-12180): at the
    recorded geometry dt equals the recorded 7.5 and the position / radius
    gradients are dL/ddt times the recorded Jacobian (radius through
    softplus_100);
  * on recorded paths of a Delaunay scene, the graph's pixel equals
    trace_one_ray's colour (the walk oracle, pinned separately) when nothing
    stops the walk early: dt from positions reproduces the recorded intervals;
  * FP64 central finite differences of the FP64 forward for log density,
    positions, SH, exposure and log radii in all three dt modes, all losses,
    ReLU and softplus, with background and opacity loss (pure relative 1e-6
    after dropping the rare coordinates whose difference quotient straddles a
    kink, detected as a disagreement between two step sizes).
"""
import math

import numpy as np
import pytest

import oracle
from bv import scenes


def tiny_params(density=4.0):
    pts = np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]], np.float64)
    return dict(log_density=np.full(4, density), positions=pts, sh=np.zeros((4, 3)), sh_degree=0)


def one_step_batch(cell, nxt, dt, origin, direction, labels):
    return dict(cells=np.array([[cell]]), next_cells=np.array([[nxt]]), dts=np.array([[dt]]), mask=np.ones((1, 1)),
                ray_origin=np.array([origin]), ray_dir=np.array([direction]), labels=np.array([labels]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_fixed_path_fixture_closed_form(dtype):
    # diff_render.rs:11961-12073 with ReLU density 4, SH 0 (colour 0.5), L1
    prm = tiny_params()
    b = one_step_batch(0, 3, 1.25, (0.1, 0.1, -1.0), (0.0, 0.0, 1.0), (0.1, 0.2, 0.3))
    r = oracle.graph_step(prm, b, loss=oracle.LOSS_L1, dtype=dtype)
    a = 1.0 - math.exp(-5.0)
    assert r["dt"][0, 0] == pytest.approx(1.25, abs=1e-6)
    assert r["pixels"][0] == pytest.approx([0.5 * a] * 3, abs=1e-6)
    assert r["loss"] == pytest.approx(sum(abs(0.5 * a - t) for t in (0.1, 0.2, 0.3)), abs=1e-6)
    gate = 0.25 / (0.25 + float(np.float32(1e-6)))
    s = 6.0 * math.exp(-5.0) * gate  # dL/dalpha = 1.5, dalpha/ddt = 4 e^-5, ddt/dt = gate
    want = np.zeros((4, 3))
    want[0] = s * np.array([0.2, 0.2, 0.5])
    want[3] = s * np.array([-0.2, -0.2, 0.5])
    np.testing.assert_allclose(r["d_positions"], want, atol=1e-6)
    # density: dL/dsigma = 1.5 dt e^-5 through ReLU (log density 4 > 0)
    assert r["d_log_density"][0] == pytest.approx(1.5 * 1.25 * math.exp(-5.0), rel=1e-5)
    assert np.all(r["d_log_density"][1:] == 0)
    # SH DC: dL/dcolour = w per channel sign, times C0
    np.testing.assert_allclose(r["d_sh"][0], [a * 0.2820948] * 3, rtol=1e-5)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_weighted_linearization_fixture(dtype):
    # diff_render.rs:12076-12180 without the oriented-surface terms: radius 1,
    # Jacobian of dt w.r.t. the current site (0.25, -0.5, 0.75, 1.25), dt 7.5
    prm = tiny_params(density=0.1)
    beta = 100.0
    prm["log_radii"] = np.full(4, math.log(math.expm1(beta * 1.0)) / beta)
    jac = np.array([0.25, -0.5, 0.75, 1.25])
    o = np.array([0.1, 0.1, -1.0])
    ref_tangent = jac[:3] @ (prm["positions"][0] - o) + jac[3] * 1.0
    b = one_step_batch(0, 3, 7.5, o, (0.0, 0.0, 1.0), (0.4, 0.6, 0.8))
    b.update(previous_cells=np.array([[0]]), dt_reference_tangent=np.array([[ref_tangent]]),
             dt_grad_previous=np.zeros((1, 1, 4)), dt_grad_current=jac.reshape(1, 1, 4),
             dt_grad_next=np.zeros((1, 1, 4)))
    r = oracle.graph_step(prm, b, loss=oracle.LOSS_L1, dt_mode=oracle.DT_LINEARIZED, dtype=dtype)
    assert r["dt"][0, 0] == pytest.approx(7.5, abs=1e-5)
    a = 1.0 - math.exp(-0.75)
    pix = 0.5 * a
    g = np.sign(pix - np.array([0.4, 0.6, 0.8]))
    dl_ddt = float(g.sum()) * 0.5 * 0.1 * math.exp(-0.75)
    want = np.zeros((4, 3))
    want[0] = dl_ddt * jac[:3]
    np.testing.assert_allclose(r["d_positions"], want, atol=1e-6)
    sig = 1.0 / (1.0 + math.exp(-beta * prm["log_radii"][0]))
    assert r["d_log_radii"][0] == pytest.approx(dl_ddt * 1.25 * sig, rel=1e-5)
    assert np.all(r["d_log_radii"][1:] == 0)


def _recorded_scene(seed=4, n=300, deg=3, rays=24, L=96):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    K = (deg + 1) ** 2
    density = np.exp(rng.normal(size=n)).astype(np.float32)
    sh = rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32)
    host = scenes.HostScene(xyz, density, sh, deg, off, nb)
    o = rng.uniform(-0.3, 0.3, size=(rays, 3)).astype(np.float32)
    d = rng.normal(size=(rays, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    starts = np.array([scenes.containing_cell(host, p) for p in o], np.uint32)
    rec = oracle.record_paths(oracle.of_host(host), o, d, starts, max_steps=L, depth=2.0, max_path_dt=50.0)
    cnt = (rec["status"] & 0x7FFFFFFF).astype(np.int64)
    mask = (np.arange(L)[None, :] < cnt[:, None]).astype(np.float32)
    cells = np.where(mask > 0, rec["cells"], 0).astype(np.uint32)
    nxt = np.where(mask > 0, rec["next_cells"], 0).astype(np.uint32)
    dts = np.where(mask > 0, rec["dts"], 0).astype(np.float32)
    return host, dict(cells=cells, next_cells=nxt, dts=dts, mask=mask, ray_origin=o, ray_dir=d,
                      labels=rng.uniform(0, 1, size=(rays, 3)).astype(np.float32)), starts


def test_graph_pixel_equals_the_walk_on_recorded_paths():
    host, b, starts = _recorded_scene()
    prm = dict(log_density=host.density, positions=host.xyz, sh=host.sh, sh_degree=host.sh_degree)
    r = oracle.graph_step(prm, b)
    tr = oracle.trace_rays(oracle.of_host(host), b["ray_origin"], b["ray_dir"], starts, weight_threshold=0.0,
                           max_steps=96, depth=2.0)
    done = (oracle.record_paths(oracle.of_host(host), b["ray_origin"], b["ray_dir"], starts, max_steps=96,
                                depth=2.0)["status"] & 0x80000000) == 0
    assert done.sum() >= 20
    np.testing.assert_allclose(r["pixels"][done], tr["rgba"][done, :3], atol=2e-5)
    np.testing.assert_allclose(r["opacity"][done], tr["rgba"][done, 3], atol=2e-5)
    np.testing.assert_allclose(r["dt"], b["dts"], atol=2e-5)


def _weighted_batch(b, rng, n):
    P, L = b["cells"].shape
    prev = np.concatenate([b["cells"][:, :1], b["cells"][:, :-1]], axis=1)
    jac = lambda: rng.normal(scale=0.3, size=(P, L, 4)) * b["mask"][..., None]
    w = dict(b)
    w.update(previous_cells=prev, dt_grad_previous=jac(), dt_grad_current=jac(), dt_grad_next=jac(),
             dt_reference_tangent=rng.normal(scale=0.01, size=(P, L)) * b["mask"])
    return w


def _fd_check(prm, b, kw, names, rng, samples=12, h=1e-4):
    """Richardson-extrapolated FP64 central differences (steps h and h/2)
    against the analytic gradient, relative 1e-6 plus the quotient's round-off
    (a few ulp of the loss over the step).  A coordinate whose two quotients
    disagree by more than 1e-3 has a kink inside the step and is skipped."""
    base = oracle.graph_step(prm, b, dtype=np.float64, **kw)
    noise = 8 * np.finfo(np.float64).eps * max(1.0, abs(base["loss"])) / h
    checked = skipped = 0
    for name in names:
        grad = base[f"d_{name}"].reshape(-1)
        flat = np.asarray(prm[name], np.float64).reshape(-1)
        nz = np.nonzero(grad)[0]
        pick = rng.choice(nz, size=min(samples, len(nz)), replace=False) if len(nz) else []
        for idx in pick:
            fd = []
            for step in (h, h / 2):
                vals = []
                for sgn in (1, -1):
                    q = dict(prm)
                    v = flat.copy()
                    v[idx] += sgn * step
                    q[name] = v.reshape(np.shape(prm[name]))
                    vals.append(oracle.graph_step(q, b, dtype=np.float64, **kw)["loss"])
                fd.append((vals[0] - vals[1]) / (2 * step))
            if abs(fd[0] - fd[1]) > 1e-3 * max(abs(fd[1]), noise):
                skipped += 1  # a kink (ReLU, clamp, |x|) inside the difference quotient
                continue
            est = (4 * fd[1] - fd[0]) / 3
            assert abs(grad[idx] - est) <= 1e-6 * max(abs(est), abs(grad[idx])) + noise, \
                (name, int(idx), grad[idx], est)
            checked += 1
    assert checked >= 3 * len(names) and skipped <= checked // 5, (checked, skipped)


@pytest.mark.parametrize("loss", [oracle.LOSS_L1, oracle.LOSS_SMOOTH_L1, oracle.LOSS_L2])
@pytest.mark.parametrize("beta", [0.0, 10.0])
def test_gradients_match_fp64_finite_differences_positions_mode(loss, beta):
    host, b, _ = _recorded_scene(seed=5, rays=12)
    rng = np.random.default_rng(7)
    # log density for ReLU is the density itself; softplus needs its inverse
    ld = host.density.astype(np.float64)
    if beta > 0:
        ld = np.log(np.expm1(beta * ld)) / beta
    prm = dict(log_density=ld, positions=host.xyz.astype(np.float64), sh=host.sh.astype(np.float64),
               sh_degree=host.sh_degree, exposure=rng.uniform(0.8, 1.2, size=(2, 3)))
    b = {k: (v.astype(np.float64) if np.asarray(v).dtype == np.float32 else v) for k, v in b.items()}
    b["view_idx"] = rng.integers(0, 2, size=b["cells"].shape[0]).astype(np.uint32)
    b["target_alpha"] = rng.uniform(0.5, 1.0, size=b["cells"].shape[0])
    kw = dict(loss=loss, softplus_beta=beta, background=(0.2, 0.0, 0.7), opacity_weight=0.3)
    _fd_check(prm, b, kw, ["log_density", "positions", "sh", "exposure"], rng)


@pytest.mark.parametrize("dt_mode", [oracle.DT_LINEARIZED, oracle.DT_RECORDED])
def test_gradients_match_fp64_finite_differences_weighted(dt_mode):
    host, b, _ = _recorded_scene(seed=6, rays=10, deg=1)
    rng = np.random.default_rng(9)
    b = {k: (v.astype(np.float64) if np.asarray(v).dtype == np.float32 else v) for k, v in b.items()}
    b = _weighted_batch(b, rng, host.n)
    prm = dict(log_density=host.density.astype(np.float64), positions=host.xyz.astype(np.float64),
               sh=host.sh.astype(np.float64), sh_degree=host.sh_degree,
               log_radii=rng.uniform(-0.05, 0.05, size=host.n))
    kw = dict(loss=oracle.LOSS_SMOOTH_L1, dt_mode=dt_mode)
    names = ["log_density", "sh"] + (["positions", "log_radii"] if dt_mode == oracle.DT_LINEARIZED else [])
    _fd_check(prm, b, kw, names, rng)
    if dt_mode == oracle.DT_RECORDED:
        r = oracle.graph_step(prm, b, dtype=np.float64, **kw)
        assert not r["d_positions"].any() and not r["d_log_radii"].any()


def test_padding_and_status_agree():
    # rows past their status count are padding: the same result with and
    # without the status hint; padded steps add nothing (diff_render.rs:11865)
    host, b, _ = _recorded_scene(seed=8, rays=8)
    prm = dict(log_density=host.density, positions=host.xyz, sh=host.sh, sh_degree=host.sh_degree)
    st = b["mask"].sum(1).astype(np.uint32)
    r0 = oracle.graph_step(prm, b, loss=oracle.LOSS_SMOOTH_L1)
    r1 = oracle.graph_step(prm, dict(b, status=st), loss=oracle.LOSS_SMOOTH_L1)
    for k in ("pixels", "dt", "d_log_density", "d_positions", "d_sh"):
        np.testing.assert_allclose(r0[k], r1[k], atol=1e-7)
    zero = dict(b, mask=np.zeros_like(b["mask"]), labels=np.zeros_like(b["labels"]))
    r = oracle.graph_step(prm, zero, loss=oracle.LOSS_SMOOTH_L1)
    assert r["loss"] == 0.0 and not r["d_positions"].any() and not r["d_sh"].any() and not r["d_log_density"].any()


def test_trainer_sh_basis_constants():
    # the trainer's basis differs from trace.rs's only in SH_C2 (1.0925484 vs
    # 1.0925485, one ulp) and in where the sign sits; its Gram matrix over the
    # sphere is the identity up to that rounding
    lib = oracle._load()
    import ctypes as C

    lib.bvo_graph_sh_basis_f32.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    rng = np.random.default_rng(0)
    d = rng.normal(size=(200_000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    B = np.zeros((len(d), 16), np.float32)
    for i in range(0, len(d), 50):
        out = np.zeros(16, np.float32)
        lib.bvo_graph_sh_basis_f32(d[i].ctypes.data, 16, out.ctypes.data)
        B[i] = out
    B = B[::50].astype(np.float64)
    gram = 4 * np.pi * (B.T @ B) / len(B)
    np.testing.assert_allclose(gram, np.eye(16), atol=0.08)
    out = np.zeros(16, np.float32)
    lib.bvo_graph_sh_basis_f32(np.array([0.6, 0.0, 0.8], np.float32).ctypes.data, 16, out.ctypes.data)
    assert out[4] == np.float32(0.0) and out[1] == np.float32(0.0) and out[3] == -np.float32(0.4886025) * np.float32(0.6)
