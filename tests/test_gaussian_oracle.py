"""The Gaussian ray-trace oracle (trace_gaussians, trace.rs:695-773) and its
adjoint, pinned without the reference running:

* the reference's closed form: two particles whose proxies overlap out of
  order composite by maximum response, rgba (0.5, 0, 0.25, 0.75)
  (trace.rs:905-933);
* the SH clamp per channel before compositing (trace.rs:784-806);
* the adversarial 65-particle fixture (tests/gaussian_gpu_vs_cpu.rs:36-74):
  the broad far particle composites after all 64 narrow ones;
* the particle frame against scipy's rotation matrices: t* and alpha equal the
  closed-form Mahalanobis minimum along the ray in FP64;
* t* is the argmax of the response, by brute-force sampling along the ray;
* the adjoint equals FP64 central finite differences of the FP64 trace, for
  every parameter kind (mean, scale, quaternion, opacity, SH).
"""
import numpy as np
import pytest

import oracle


def dc(c):
    return (np.asarray(c, np.float64) - 0.5) / 0.2820948


def sh1(color, z_delta):
    out = np.zeros(12)
    out[:3] = dc(color)
    out[6:9] = np.asarray(z_delta, np.float64) / 0.48860252
    return out


def two_gaussians():
    # index order far, near: the far broad particle's support begins first
    return oracle.GaussianScene(np.array([[0, 0, 4], [0, 0, 2]], np.float32), np.array([0.5, 0.5], np.float32),
                                np.array([[0, 0, 0, 1], [0, 0, 0, 1]], np.float32),
                                np.array([[1, 1, 1], [1, 1, 0.1]], np.float32),
                                np.stack([dc([0, 0, 1]), dc([1, 0, 0])]).astype(np.float32), 0)


def adversarial():
    """tests/gaussian_gpu_vs_cpu.rs:36-74."""
    mean = [[0, 0, 8]] + [[0, 0, 2 + 5 * i / 63] for i in range(64)]
    op = [0.15] + [0.015] * 64
    sh = [sh1([0, 0, 0.8], [0, 0.1, 0.2])]
    for i in range(64):
        t = i / 63
        sh.append(sh1([0.8 * (1 - t), 0.2 * t, 0], [0.2 * (1 - t), 0.1, 0.2 * t]))
    scale = [[1, 1, 7]] + [[1, 1, 0.02]] * 64
    return oracle.GaussianScene(np.array(mean, np.float32), np.array(op, np.float32),
                                np.tile(np.array([0, 0, 0, 1], np.float32), (65, 1)), np.array(scale, np.float32),
                                np.array(sh, np.float32), 1)


def test_two_particles_closed_form():
    r = oracle.trace_gaussians(two_gaussians(), [[0, 0, 0]], [[0, 0, 1]], min_opacity=0.01, min_transmittance=0.0,
                               t_end=10.0, record=4)
    np.testing.assert_allclose(r["rgba"][0], [0.5, 0.0, 0.25, 0.75], atol=1e-6)
    assert r["hits"][0] == 2 and list(r["order"][0, :2]) == [1, 0]


def test_sh_clamp_before_compositing():
    g = oracle.GaussianScene(np.array([[0, 0, 1]], np.float32), np.array([0.999], np.float32),
                             np.array([[0, 0, 0, 1]], np.float32), np.array([[1, 1, 1]], np.float32),
                             dc([-1.0, 0.25, 2.0])[None].astype(np.float32), 0)
    r = oracle.trace_gaussians(g, [[0, 0, 0]], [[0, 0, 1]], min_opacity=0.01, min_transmittance=0.0, t_end=10.0)
    a = 0.999
    np.testing.assert_allclose(r["rgba"][0], [0.0, 0.25 * a, 2.0 * a, a], atol=2e-6)


def test_adversarial_fixture_orders_broad_particle_last():
    g = adversarial()
    r = oracle.trace_gaussians(g, [[0, 0, 0]], [[0, 0, 1]], min_opacity=0.01, min_transmittance=0.0, t_end=10.0,
                               record=80)
    assert r["hits"][0] == 65
    assert list(r["order"][0, :65]) == list(range(1, 65)) + [0]


def _quat_matrix(q):
    from scipy.spatial.transform import Rotation

    return Rotation.from_quat(np.asarray(q, np.float64)).as_matrix()


def random_gaussians(rng, n, deg=1, spread=0.6, center=(0, 0, 3)):
    mean = np.asarray(center) + rng.uniform(-spread, spread, (n, 3))
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scale = np.exp(rng.uniform(np.log(0.15), np.log(0.6), (n, 3)))
    op = rng.uniform(0.3, 0.9, n)
    K = (deg + 1) ** 2
    sh = rng.normal(scale=0.3, size=(n, 3 * K))
    sh[:, :3] += 1.0  # keep the colour away from the clamp
    return oracle.GaussianScene(mean, op, q, scale, sh, deg)


def test_frame_matches_rotation_matrices_fp64():
    rng = np.random.default_rng(0)
    g = random_gaussians(rng, 1, spread=0.2)
    o = np.zeros(3)
    d = np.array([0.05, -0.03, 1.0])
    r = oracle.trace_gaussians(g, o[None], d[None], min_opacity=1e-9, min_transmittance=0.0, t_end=100.0,
                               dtype=np.float64)
    dn = d / np.linalg.norm(d)
    R = _quat_matrix(g.rotation[0])
    Sinv = R @ np.diag(1 / np.asarray(g.scale[0]) ** 2) @ R.T
    u = o - g.mean[0]
    t = -(u @ Sinv @ dn) / (dn @ Sinv @ dn)
    q = (u + t * dn) @ Sinv @ (u + t * dn)
    alpha = g.opacity[0] * np.exp(-0.5 * q)
    assert abs(r["rgba"][0, 3] - alpha) < 1e-12


def test_max_response_is_argmax_brute_force():
    rng = np.random.default_rng(1)
    g = random_gaussians(rng, 1, spread=0.3)
    for k in range(8):
        o = rng.uniform(-0.3, 0.3, 3)
        d = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), 1.0])
        dn = d / np.linalg.norm(d)
        R = _quat_matrix(g.rotation[0])
        Sinv = R @ np.diag(1 / np.asarray(g.scale[0]) ** 2) @ R.T
        ts = np.linspace(0, 8, 80001)
        x = o[None] + ts[:, None] * dn - g.mean[0]
        resp = np.exp(-0.5 * np.einsum("ij,jk,ik->i", x, Sinv, x))
        r = oracle.trace_gaussians(g, o[None], d[None], min_opacity=1e-12, min_transmittance=0.0, t_end=100.0,
                                   dtype=np.float64)
        assert abs(r["rgba"][0, 3] - g.opacity[0] * resp.max()) < 1e-6


def _loss(g, o, d, w):
    r = oracle.trace_gaussians(g, o, d, min_opacity=0.01, min_transmittance=0.0, t_end=100.0, dtype=np.float64)
    return float((r["rgba"] * w).sum())


@pytest.mark.parametrize("seed", [0, 1])
def test_adjoint_matches_fp64_finite_differences(seed):
    rng = np.random.default_rng(seed)
    g = random_gaussians(rng, 6, deg=2, spread=0.35)
    m = 5
    o = np.tile([0.0, 0.0, 0.0], (m, 1))
    d = np.concatenate([rng.uniform(-0.08, 0.08, (m, 2)), np.ones((m, 1))], 1)
    w = rng.normal(size=(m, 4))
    adj = oracle.adjoint_gaussians(g, o, d, w, min_opacity=0.01, min_transmittance=0.0, t_end=100.0,
                                   dtype=np.float64)
    checked = 0
    for field in ("mean", "scale", "rotation", "opacity", "sh"):
        arr = getattr(g, field)
        flat = arr.reshape(-1)
        for j in rng.choice(flat.size, size=min(flat.size, 10), replace=False):
            h = 1e-6 * max(1.0, abs(flat[j]))
            keep = flat[j]
            flat[j] = keep + h
            lp = _loss(g, o, d, w)
            flat[j] = keep - h
            lm = _loss(g, o, d, w)
            flat[j] = keep
            fd = (lp - lm) / (2 * h)
            an = adj[field].reshape(-1)[j]
            assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (field, j, an, fd)
            checked += 1
    assert checked >= 40
    # the FP32 adjoint agrees with the FP64 one
    g32 = oracle.GaussianScene(*(np.asarray(getattr(g, f), np.float32) for f in
                                 ("mean", "opacity", "rotation", "scale", "sh")), g.sh_degree)
    a32 = oracle.adjoint_gaussians(g32, o, d, w, min_opacity=0.01, min_transmittance=0.0, t_end=100.0)
    for field in ("mean", "scale", "rotation", "opacity", "sh"):
        ref = adj[field]
        assert np.abs(a32[field] - ref).max() <= 1e-3 * max(1.0, np.abs(ref).max()), field
