"""The PowerFoam splat oracle (trace_powerfoam_splats, trace.rs:1331-1546) pinned
without the reference running:

* the reference's own closed form: two disconnected supports are both found,
  in depth order, with dt = 1 each (trace.rs:2397-2442);
* a single support: the chord of a sphere, alpha = 1 - exp(-sigma * 2 r);
* the camera exclusion (no interval for a support within 4 r of the camera);
* the power-diagram invariant, by brute force over every site: inside the union
  of the support balls, the point at the middle of each interval has the
  interval's cell as its power-nearest site, the intervals are sorted and
  disjoint, and together they cover every sampled point that lies inside some
  ball (a dropped plane, a wrong shift sign or a wrong sphere root fails these);
* the Cech builder against an O(N^2) brute force;
* trace and record agree: opacity = 1 - prod(1 - alpha_i) over recorded rows.
"""
import numpy as np
import pytest

import oracle
from bv import scenes


def _scene(xyz, radius, density=None, sh=None, deg=0, adjacency=None):
    xyz = np.asarray(xyz, np.float32)
    n = len(xyz)
    radius = np.asarray(radius, np.float32)
    if adjacency is None:
        off, nb = scenes.cech_csr(xyz, radius)
    else:
        off, nb = adjacency
    K = (deg + 1) ** 2
    return oracle.Scene(xyz, np.ones(n, np.float32) if density is None else np.asarray(density, np.float32),
                        np.zeros((n, 3 * K), np.float32) if sh is None else np.asarray(sh, np.float32), deg, off, nb,
                        radius)


def test_disconnected_supports_in_depth_order():
    # trace.rs:2397-2442: spheres at z = 3 and 6 of radius 0.5, empty adjacency
    sc = _scene([[0, 0, 3], [0, 0, 6]], [0.5, 0.5], adjacency=(np.zeros(3, np.uint32), np.zeros(0, np.uint32)))
    r = oracle.record_splats(sc, [[0, 0, 0]], [[0, 0, 1]], max_steps=8, depth=10.0)
    assert r["count"][0] == 2
    assert list(r["cells"][0, :2]) == [0, 1]
    assert np.all(np.abs(r["dts"][0, :2] - 1.0) < 1e-6)
    t = oracle.trace_splats(sc, [[0, 0, 0]], [[0, 0, 1]], depth=10.0)
    a = 1 - np.exp(-1.0)
    assert abs(t["rgba"][0, 3] - (1 - (1 - a) ** 2)) < 1e-6
    assert t["steps"][0] == 2 and t["last_point"][0] == 1 and abs(t["t_end"][0] - 6.5) < 1e-6


def test_single_support_chord_and_colour():
    # the batched-rows fixture of tests/radfoam_gpu_vs_cpu.rs:700-713: one site at
    # z = 5, radius 1.2 (SH DC 0.4, -0.2, 0.1)
    sh = np.array([[0.4, -0.2, 0.1]], np.float32)
    sc = _scene([[0, 0, 5]], [1.2], density=[0.7], sh=sh, adjacency=(np.zeros(2, np.uint32), np.zeros(0, np.uint32)))
    t = oracle.trace_splats(sc, [[0, 0, 0]], [[0, 0, 2.0]], depth=10.0, weight_threshold=1e-4, max_steps=4)
    alpha = 1 - np.exp(-0.7 * 2.4)
    colour = np.maximum(0.5 + 0.2820948 * sh[0], 0)
    np.testing.assert_allclose(t["rgba"][0], np.r_[alpha * colour, alpha], atol=2e-6)
    assert abs(t["depth_mode"][0] - 5.0) < 1e-5 and abs(t["t_end"][0] - 6.2) < 1e-5
    # a ray that misses and a ray whose depth ends before the sphere
    miss = oracle.trace_splats(sc, [[2, 0, 0]], [[0, 0, 1]], depth=10.0)
    assert miss["steps"][0] == 0 and miss["rgba"][0, 3] == 0
    short = oracle.trace_splats(sc, [[0, 0, 0]], [[0, 0, 1]], depth=3.7)
    assert short["steps"][0] == 0


def test_camera_exclusion():
    # a support whose centre is within 4 r of the camera is skipped
    sc = _scene([[0, 0, 1.0], [0, 0, 5.0]], [0.3, 0.5], adjacency=(np.zeros(3, np.uint32), np.zeros(0, np.uint32)))
    r = oracle.record_splats(sc, [[0, 0, 0]], [[0, 0, 1]], max_steps=4, depth=10.0)
    assert r["count"][0] == 1 and r["cells"][0, 0] == 1
    sc2 = _scene([[0, 0, 1.3], [0, 0, 5.0]], [0.3, 0.5], adjacency=(np.zeros(3, np.uint32), np.zeros(0, np.uint32)))
    r2 = oracle.record_splats(sc2, [[0, 0, 0]], [[0, 0, 1]], max_steps=4, depth=10.0)
    assert r2["count"][0] == 2 and list(r2["cells"][0, :2]) == [0, 1]


def test_cech_builder_brute_force():
    rng = np.random.default_rng(5)
    n = 1500
    xyz = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(1e-3), np.log(0.3), n)).astype(np.float32)
    off, nb = scenes.cech_csr(xyz, r)
    d = xyz[:, None, :] - xyz[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    s = r[:, None] + r[None, :]
    A = d2 <= s * s
    np.fill_diagonal(A, False)
    for i in range(n):
        np.testing.assert_array_equal(nb[off[i]:off[i + 1]], np.nonzero(A[i])[0])


def _power_argmin(xyz, r, pts):
    xyz64, r64 = xyz.astype(np.float64), r.astype(np.float64)
    pw = ((pts[:, None, :] - xyz64[None]) ** 2).sum(-1) - r64[None] ** 2
    return np.argmin(pw, 1), pw.min(1)


@pytest.mark.parametrize("seed", [0, 1])
def test_power_diagram_invariants(seed):
    rng = np.random.default_rng(seed)
    n = 600
    xyz = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(0.05), np.log(0.35), n)).astype(np.float32)
    sc = _scene(xyz, r)
    m = 40
    o = np.tile(np.array([[0.0, 0.0, 4.0]], np.float32), (m, 1))
    tgt = rng.uniform(-0.8, 0.8, (m, 3))
    d = (tgt - o).astype(np.float32)
    rec = oracle.record_splats(sc, o, d, max_steps=4096, depth=100.0)
    for k in range(m):
        cnt = int(rec["count"][k])
        cells, nears, dts = rec["cells"][k, :cnt], rec["nears"][k, :cnt].astype(np.float64), rec["dts"][k, :cnt]
        assert cnt < 4096
        # sorted by (near, cell) and disjoint
        assert np.all(np.diff(nears) >= 0)
        assert np.all(nears[1:] >= nears[:-1] + dts[:-1] - 1e-5)
        dirn = d[k].astype(np.float64) / np.linalg.norm(d[k].astype(np.float64))
        # interval midpoints belong to their cell (skip slivers)
        mids = nears + 0.5 * dts
        keep = dts > 1e-4
        pts = o[k].astype(np.float64) + mids[keep, None] * dirn
        arg, pmin = _power_argmin(xyz, r, pts)
        np.testing.assert_array_equal(arg, cells[keep])
        assert np.all(pmin <= 1e-6)
        # coverage: a sample inside some ball lies in one of the intervals
        ts = np.linspace(0.0, 10.0, 4001)
        pts = o[k].astype(np.float64) + ts[:, None] * dirn
        _, pm = _power_argmin(xyz, r, pts)
        inside = pm < -1e-4
        covered = np.zeros_like(inside)
        for a, b in zip(nears, nears + dts):
            covered |= (ts >= a - 1e-6) & (ts <= b + 1e-6)
        assert np.all(covered[inside])
        # and nothing outside every ball is covered (beyond rounding)
        assert not np.any(covered & (pm > 1e-4))


def test_trace_matches_record_and_truncation():
    rng = np.random.default_rng(3)
    sc_h = scenes.powerfoam_scene(4000, seed=3, sh_degree=1, density_scale=2.0)
    sc = oracle.of_host(sc_h)
    m = 64
    o = np.tile(np.array([[0.0, 0.2, 3.0]], np.float32), (m, 1))
    d = (rng.uniform(-0.6, 0.6, (m, 3)) - o).astype(np.float32)
    rec = oracle.record_splats(sc, o, d, max_steps=1024, depth=100.0)
    tr = oracle.trace_splats(sc, o, d, depth=100.0, weight_threshold=0.0, eval_mode=oracle.EVAL_OPACITY)
    for k in range(m):
        cnt = int(rec["count"][k])
        dens = sc.density[rec["cells"][k, :cnt]].astype(np.float64)
        alpha = np.where(dens > 1e-6, 1 - np.exp(-dens * rec["dts"][k, :cnt]), 0.0)
        assert abs((1 - np.prod(1 - alpha)) - tr["rgba"][k, 3]) < 1e-5
        assert tr["steps"][k] == cnt
    assert rec["count"].max() > 16  # rows longer than the device k-buffer
    # truncation to max_steps keeps the first intervals
    tr3 = oracle.trace_splats(sc, o, d, depth=100.0, weight_threshold=0.0, max_steps=3)
    assert np.all(tr3["steps"] == np.minimum(rec["count"], 3))


def test_splat_relight_single_support_closed_form():
    """One support at z = 5 (radius 1, density s), camera at the origin along +z,
    one isotropic light at (0, 0, -3), diffuse only: the interval starts at
    x = (0, 0, 4) with normal (0, 0, -1) facing the light, so the colour is
    albedo * I / |x - light|^2 * (1 - exp(-2 s)); the shadow trace from the
    cell's depth point excludes the cell itself (4 r camera exclusion), so
    visibility changes nothing."""
    s = 0.8
    sc = _scene([[0, 0, 5]], [1.0], density=[s], adjacency=(np.zeros(2, np.uint32), np.zeros(0, np.uint32)))
    mat = np.array([[0.3, 0.6, 0.9, 0.5]], np.float32)
    light = dict(position=np.array([0, 0, -3], np.float32), intensity=np.array([4.0, 2.0, 1.0], np.float32),
                 exponent=0.0)
    alpha = 1 - np.exp(-2 * s)
    want = mat[0, :3] * light["intensity"] / 49.0 * alpha
    for vis in (False, True):
        got = oracle.trace_splat_relight(sc, mat, [[0, 0, 0]], [[0, 0, 1]], [light], visibility=vis,
                                         specular_f0=(0, 0, 0), depth=10.0)
        np.testing.assert_allclose(got[0, :3], want, rtol=2e-6, atol=1e-7)
        assert abs(got[0, 3] - alpha) < 1e-6


def test_splat_relight_occluder_scales_light():
    """An occluding support between the depth point and the light scales that
    light's contribution by its transmittance exp(-s_o * chord)."""
    so = 1.5
    # lit cell at z = 5 (radius 1); the light at (0, 3, 1) faces the entry point
    # (0, 0, 4); the shadow ray runs from the depth point (0, 0, 5) to the light
    # and passes through the centre of an occluder of radius 0.4 at (0, 1.5, 3)
    # (off the camera ray), a chord of 0.8
    sc = _scene([[0, 0, 5], [0, 1.5, 3]], [1.0, 0.4], density=[0.8, so],
                adjacency=(np.zeros(3, np.uint32), np.zeros(0, np.uint32)))
    mat = np.array([[0.3, 0.6, 0.9, 0.5], [1, 1, 1, 0.5]], np.float32)
    light = dict(position=np.array([0, 3, 1], np.float32), intensity=np.array([4.0, 2.0, 1.0], np.float32),
                 exponent=0.0)
    free = oracle.trace_splat_relight(sc, mat, [[0, 0, 0]], [[0, 0, 1]], [light], visibility=False,
                                      specular_f0=(0, 0, 0), depth=10.0)
    shad = oracle.trace_splat_relight(sc, mat, [[0, 0, 0]], [[0, 0, 1]], [light], visibility=True,
                                      specular_f0=(0, 0, 0), depth=10.0)
    assert free[0, :3].max() > 0
    np.testing.assert_allclose(shad[0, :3], free[0, :3] * np.exp(-so * 0.8), rtol=1e-5)
    assert shad[0, 3] == free[0, 3]
