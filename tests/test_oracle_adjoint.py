"""The oracle's analytic adjoint (FP64) against central finite differences of
its own FP64 forward.  This pins the gradient formulas the HIP adjoint is then
compared with.  CPU only."""
import numpy as np

import oracle


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.3, 0.3, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o, d = o.astype(np.float32), d.astype(np.float32)
    from bv import scenes

    starts = np.array([scenes.containing_cell(scene, x) for x in o], np.uint32)
    return o, d, starts


def _loss(scene64, o, d, s, target, lam, wt):
    r = oracle.trace_rays(scene64, o, d, s, weight_threshold=wt, max_steps=1024, depth=3.0, dtype=np.float64)
    e = r["rgba"][:, :3] - target
    return float((e * e).sum() + (lam * r["rgba"][:, 3]).sum()), r


def test_adjoint_f64_matches_central_differences(small_scene):
    sc = small_scene
    o, d, s = _rays(sc, 48, 1)
    rng = np.random.default_rng(2)
    target = rng.uniform(0, 1, size=(48, 3))
    lam = rng.normal(size=48)
    wt = 0.001
    s64 = oracle.of_host(sc).as_dtype(np.float64)
    L0, r0 = _loss(s64, o, d, s, target, lam, wt)
    g = np.concatenate([2 * (r0["rgba"][:, :3] - target), lam[:, None]], axis=1)
    adj = oracle.adjoint(s64, o, d, s, g, weight_threshold=wt, max_steps=1024, depth=3.0, dtype=np.float64)
    rec = oracle.record_paths(oracle.of_host(sc), o, d, s, max_steps=256, depth=3.0)
    touched = np.unique(rec["cells"][rec["cells"] != oracle.NO_CELL])
    picks = rng.choice(touched, size=min(12, len(touched)), replace=False)
    h = 1e-6
    checked = 0
    for cell in picks:
        for kind, idx in (("density", (cell,)), ("sh", (cell, int(rng.integers(0, 48)))), ("sh", (cell, 0)),
                          ("xyz", (cell, int(rng.integers(0, 3))))):
            arr = getattr(s64, kind)
            base = arr[idx]
            arr[idx] = base + h
            lp, _ = _loss(s64, o, d, s, target, lam, wt)
            arr[idx] = base - h
            lm, _ = _loss(s64, o, d, s, target, lam, wt)
            arr[idx] = base
            fd = (lp - lm) / (2 * h)
            an = adj[kind][idx]
            scale = max(abs(fd), 1e-3)
            assert abs(an - fd) <= 1e-5 * scale + 1e-7, (kind, idx, an, fd)
            checked += 1
    assert checked >= 24


def test_adjoint_f32_close_to_f64(small_scene):
    sc = small_scene
    o, d, s = _rays(sc, 64, 3)
    rng = np.random.default_rng(4)
    g = rng.normal(size=(64, 4))
    a32 = oracle.adjoint(oracle.of_host(sc), o, d, s, g, depth=3.0)
    a64 = oracle.adjoint(oracle.of_host(sc).as_dtype(np.float64), o, d, s, g, depth=3.0, dtype=np.float64)
    for k in ("density", "sh", "xyz"):
        ref = a64[k]
        err = np.abs(a32[k] - ref).max()
        assert err <= 1e-4 * max(np.abs(ref).max(), 1.0), (k, err)


def test_threaded_adjoint_equals_serial_up_to_summation_order(small_scene):
    """The CPU baseline's threaded FP32 adjoint (per-thread ray ranges, atomic
    adds) against the serial oracle: the same per-ray terms summed in another
    order, so within FP32 summation-order error of the largest gradient."""
    sc = oracle.of_host(small_scene)
    o, d, s = _rays(small_scene, 256, 7)
    g = np.random.default_rng(8).normal(size=(256, 4))
    ser = oracle.adjoint(sc, o, d, s, g, depth=3.0)
    for threads in (1, 4, 0):
        mt = oracle.adjoint(sc, o, d, s, g, depth=3.0, threads=threads)
        for k in ("density", "sh", "xyz"):
            err = np.abs(mt[k] - ser[k]).max()
            assert err <= 1e-5 * max(np.abs(ser[k]).max(), 1.0), (threads, k, err)
        assert np.any(mt["sh"] != 0)
