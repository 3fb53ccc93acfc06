"""Pins the CPU oracle (oracle/bv_oracle.cpp) to the reference's own closed-form
tests and to invariants that a dropped term, a wrong sign or a wrong index would
break.  CPU only."""
import math
import os

import numpy as np
import pytest

import oracle
from bv import io, scenes
from conftest import GOLDEN
import fixtures


def one_cell(density, radius=None):
    return oracle.Scene(np.zeros((1, 3), np.float32), np.array([density], np.float32), np.zeros((1, 3), np.float32),
                        0, np.array([0, 0], np.uint32), np.zeros(0, np.uint32),
                        None if radius is None else np.array([radius], np.float32))


def test_terminal_cell_integrates_to_depth():
    # trace.rs:2210-2244: one unit cell, no exit face, density 2, depth 3 -> dt 3, alpha 1 - e^-6.
    sc = one_cell(2.0)
    rec = oracle.record_paths(sc, [[0, 0, 0]], [[0, 0, 1]], 0, max_steps=32, depth=3.0)
    assert rec["status"][0] == 1 and rec["cells"][0, 0] == 0 and rec["dts"][0, 0] == 3.0
    r = oracle.trace_rays(sc, [[0, 0, 0]], [[0, 0, 1]], 0, weight_threshold=0.0, max_steps=32, depth=3.0,
                          eval_mode=oracle.EVAL_CONSTANT)
    assert abs(r["rgba"][0, 3] - (1 - np.exp(np.float32(-6.0)))) < 1e-6
    assert r["t_end"][0] == 3.0


def test_power_foam_clips_to_support_sphere():
    # trace.rs:2246-2280: radius 1 sphere at the origin, ray from z=-2: dt = 2, alpha = 1 - e^-4.
    sc = one_cell(2.0, radius=1.0)
    r = oracle.trace_rays(sc, [[0, 0, -2]], [[0, 0, 1]], 0, weight_threshold=0.0, max_steps=32, depth=10.0,
                          eval_mode=oracle.EVAL_CONSTANT)
    assert abs(r["rgba"][0, 3] - (1 - math.exp(-4.0))) < 1e-6
    assert r["t_end"][0] == 10.0
    assert abs(r["depth_mode"][0] - 2.0) < 1e-6  # mid-segment of [1, 3]


def test_support_interval_keeps_non_unit_parameterisation():
    # trace.rs:2444-2461.
    o = [0, 0, -2]
    assert oracle.support_interval(True, o, [0, 0, 1], [0, 0, 0], 1.0, 0.0, 10.0) == (1.0, 3.0)
    assert oracle.support_interval(True, o, [0, 0, 2], [0, 0, 0], 1.0, 0.0, 10.0) == (0.5, 1.5)
    assert oracle.support_interval(False, o, [0, 0, 1], [0, 0, 0], 1.0, 0.0, 10.0) == (0.0, 10.0)
    assert oracle.support_interval(True, o, [0, 1, 0], [0, 0, 0], 1.0, 0.0, 10.0) is None


def brute_force_intervals(xyz, radius, ro, rd, depth):
    """O(N^2) cell intervals along a unit ray: each site's power cell intersected
    with every radical half-space (trace.rs:2463-2549), in float64."""
    n = xyz.shape[0]
    out = []
    for i in range(n):
        near, far = 0.0, depth
        for j in range(n):
            if i == j:
                continue
            off = xyz[j] - xyz[i]
            ri = 0.0 if radius is None else radius[i]
            rj = 0.0 if radius is None else radius[j]
            shift = 0.5 + 0.5 * (ri * ri - rj * rj) / off.dot(off)
            face = xyz[i] + shift * off
            den = off.dot(rd)
            t = (face - ro).dot(off) / den if den != 0 else None
            if den > 0:
                far = min(far, t)
            elif den < 0:
                near = max(near, t)
        if radius is not None:
            oc = ro - xyz[i]
            b = oc.dot(rd)
            disc = b * b - (oc.dot(oc) - radius[i] ** 2)
            if disc <= 0:
                continue
            near, far = max(near, -b - math.sqrt(disc)), min(far, -b + math.sqrt(disc))
        if far > near:
            out.append((near, i, far - near))
    out.sort()
    return out


def test_power_foam_walk_matches_brute_force_three_sites():
    # trace.rs:2463-2549 restated.
    xyz = np.array([[0, 0, 0], [2, 0, 0], [4, 0, 0]], np.float32)
    rad = np.array([1.4, 0.9, 1.6], np.float32)
    off, nb = scenes.csr_from_rows([[1, 2], [0, 2], [0, 1]])
    sc = oracle.Scene(xyz, np.ones(3, np.float32), np.zeros((3, 3), np.float32), 0, off, nb, rad)
    walked = oracle.record_paths(sc, [[-2, 0, 0]], [[1, 0, 0]], 0, max_steps=32, depth=8.0)
    bf = brute_force_intervals(xyz.astype(np.float64), rad.astype(np.float64), np.array([-2.0, 0, 0]),
                               np.array([1.0, 0, 0]), 8.0)
    # The recorder emits every cell interval (support clipping only changes dt,
    # which the trace applies); compare the cell order of the clipped brute force
    # with the walk's cells that intersect their support.
    r = oracle.trace_rays(sc, [[-2, 0, 0]], [[1, 0, 0]], 0, weight_threshold=0.0, max_steps=32, depth=8.0,
                          eval_mode=oracle.EVAL_OPACITY)
    count = walked["status"][0] & 0x7FFFFFFF
    cells = list(walked["cells"][0, :count])
    assert [c for _, c, _ in bf] == [c for c in cells if c in {i for _, i, _ in bf}]
    # alpha from the brute-force clipped intervals equals the traced alpha.
    T = 1.0
    for _, i, dt in bf:
        T *= math.exp(-1.0 * dt)
    assert abs((1 - T) - r["rgba"][0, 3]) < 1e-5


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_voronoi_walk_matches_brute_force_on_delaunay(seed):
    """On a random Delaunay scene the walk visits exactly the cells whose
    brute-force Voronoi interval is non-empty, in order, with the same lengths."""
    rng = np.random.default_rng(100 + seed)
    xyz = rng.uniform(-1, 1, size=(300, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    sc = oracle.Scene(xyz, np.ones(300, np.float32), np.zeros((300, 3), np.float32), 0, off, nb)
    for k in range(6):
        ro = rng.uniform(-0.5, 0.5, 3).astype(np.float32)
        rd = rng.normal(size=3)
        rd = (rd / np.linalg.norm(rd)).astype(np.float32)
        start = scenes.containing_cell(scenes.HostScene(xyz, sc.density, sc.sh, 0, off, nb), ro)
        rec = oracle.record_paths(sc, ro[None], rd[None], start, max_steps=512, depth=5.0)
        cnt = rec["status"][0] & 0x7FFFFFFF
        bf = brute_force_intervals(xyz.astype(np.float64), None, ro.astype(np.float64),
                                   rd.astype(np.float64) / np.linalg.norm(rd.astype(np.float64)), 5.0)
        bf = [b for b in bf if b[2] > 1e-6]  # drop slivers below FP32 resolution
        walk = [(int(c), float(dt)) for c, dt in zip(rec["cells"][0, :cnt], rec["dts"][0, :cnt]) if dt > 1e-6]
        assert [c for c, _ in walk] == [i for _, i, _ in bf]
        np.testing.assert_allclose([dt for _, dt in walk], [b[2] for b in bf], atol=2e-5)


def test_record_path_alpha_equals_trace_alpha():
    # trace.rs:2164-2194 on the tetra fixture.
    t = oracle.of_host(fixtures.tetra())
    ro, rd = [[0.1, 0.1, -1.0]], [[0, 0, 1]]
    rec = oracle.record_paths(t, ro, rd, 0, max_steps=32, depth=100.0)
    tr = oracle.trace_rays(t, ro, rd, 0, weight_threshold=0.0, max_steps=32, depth=100.0,
                           eval_mode=oracle.EVAL_CONSTANT)
    cnt = rec["status"][0] & 0x7FFFFFFF
    T, acc = np.float32(1), np.float32(0)
    for c, dt in zip(rec["cells"][0, :cnt], rec["dts"][0, :cnt]):
        a = np.float32(1) - np.exp(-np.float32(5.0) * dt)
        acc += T * a
        T *= np.float32(1) - a
    assert abs(acc - tr["rgba"][0, 3]) < 1e-4
    rec2 = oracle.record_paths(t, ro, rd, 0, max_steps=2, depth=100.0)
    assert (rec2["status"][0] & 0x7FFFFFFF) <= 2


def test_recorder_truncation_bit_and_dt_clamp():
    # radfoam_record_paths.wgsl:1339-1354: bit 31 set when the row fills with a face still ahead.
    g = oracle.of_host(fixtures.x_grid(16))
    o, d = [[-0.25, 0.0, 0.0]], [[1.0, 0.0, 0.0]]
    full = oracle.record_paths(g, o, d, 0, max_steps=64, depth=100.0)
    assert full["status"][0] == 16
    np.testing.assert_array_equal(full["cells"][0, :16], np.arange(16))
    np.testing.assert_array_equal(full["next_cells"][0, :15], np.arange(1, 16))
    assert full["next_cells"][0, 15] == 15  # terminal: next == cell
    np.testing.assert_allclose(full["dts"][0, 1:15], 1.0)
    assert abs(full["dts"][0, 0] - 0.75) < 1e-6
    trunc = oracle.record_paths(g, o, d, 0, max_steps=4, depth=100.0)
    assert trunc["status"][0] == (4 | oracle.TRUNCATED_BIT)
    clamped = oracle.record_paths(g, o, d, 0, max_steps=64, depth=100.0, max_path_dt=0.5)
    assert np.all(clamped["dts"][0, :16] <= 0.5)
    assert clamped["dts"][0, 15] == 0.5  # last segment 15 -> 100 clamped


def test_tiny_ply_fixture_closed_form():
    # tests/radfoam_cpu_ref.rs:17-42 on tests/data/radfoam_tiny_ascii.ply: four
    # coplanar sites, a ray along +z has no face with o.d > 0, so cell 0 absorbs
    # until depth: alpha = 1 - exp(-0.1 * 100).
    m = io.load_radfoam_ascii(os.path.join(GOLDEN, "radfoam_tiny_ascii.ply"))
    assert m.n == 4 and m.sh_degree == 1 and list(m.offsets) == [0, 3, 6, 9, 12]
    expected_dc = ((128.0 / 255.0) - 0.5) / 0.2820948
    assert np.all(np.abs(m.sh[:, :3] - expected_dc) <= 1e-3)
    r = oracle.trace_rays(oracle.of_host(m), [[0.5, 0.5, -1.0]], [[0, 0, 1]], 0, weight_threshold=1e-4,
                          max_steps=16, depth=100.0, eval_mode=oracle.EVAL_CONSTANT)
    a = 1 - np.exp(np.float32(-0.1) * np.float32(100.0))
    np.testing.assert_allclose(r["rgba"][0], [a, a, a, a], atol=1e-6)
    assert r["steps"][0] == 1


def test_chain_fixture_closed_form():
    # tests/radfoam_synth_chain.rs: sites on z at spacing dz; a +z ray through all
    # cells integrates density * total length.
    ch = fixtures.chain(5, 1.0, 0.1, 1)
    assert list(ch.offsets) == [0, 1, 3, 5, 7, 8]
    r = oracle.trace_rays(oracle.of_host(ch), [[0, 0, -0.5]], [[0, 0, 1]], 0, weight_threshold=0.0, max_steps=64,
                          depth=10.0, eval_mode=oracle.EVAL_CONSTANT)
    assert abs(r["rgba"][0, 3] - (1 - math.exp(-0.1 * 10.0))) < 1e-5
    assert r["last_point"][0] == 4


def test_sh_clamps_per_cell_before_compositing():
    # trace.rs:784-806.
    dc = (np.array([-1.0, 0.25, 2.0], np.float32) - np.float32(0.5)) / np.float32(0.2820948)
    sc = oracle.Scene(np.zeros((1, 3), np.float32), np.ones(1, np.float32), dc[None], 0, np.array([0, 0], np.uint32),
                      np.zeros(0, np.uint32))
    c = oracle.eval_sh(sc, 0, [0, 0, 1])
    assert c[0] == 0.0 and abs(c[1] - 0.25) < 1e-6 and abs(c[2] - 2.0) < 1e-6


def test_sh_basis_is_orthonormal():
    """The 16 constants times their polynomials are an orthonormal basis of the
    real SH up to degree 3: a wrong constant, factor or cross term fails the
    Gram matrix.  Coefficient k alone isolates basis k through eval_sh."""
    K = 16
    nodes, weights = np.polynomial.legendre.leggauss(12)
    phis = np.linspace(0, 2 * np.pi, 24, endpoint=False)
    dirs, w = [], []
    for ct, wt in zip(nodes, weights):
        st = math.sqrt(1 - ct * ct)
        for ph in phis:
            dirs.append([st * math.cos(ph), st * math.sin(ph), ct])
            w.append(wt * 2 * np.pi / len(phis))
    dirs = np.asarray(dirs, np.float32)
    # A coefficient of 0.25 keeps 0.5 + 0.25 b inside (0, 1), so the clamp never hides the basis value.
    B2 = np.zeros((len(dirs), K))
    for k in range(K):
        sh = np.zeros((1, 48), np.float32)
        sh[0, 3 * k] = 0.25
        sc = oracle.Scene(np.zeros((1, 3), np.float32), np.ones(1, np.float32), sh, 3, np.array([0, 0], np.uint32),
                          np.zeros(0, np.uint32))
        B2[:, k] = (np.array([oracle.eval_sh(sc, 0, d)[0] for d in dirs], np.float64) - 0.5) / 0.25
    G = (B2 * np.asarray(w)[:, None]).T @ B2
    np.testing.assert_allclose(G, np.eye(K), atol=2e-5)


def test_camera_ray_matches_quaternion_rotation():
    cam = scenes.looking_at((2.0, 1.5, -3.0), (0.0, 0.5, 0.0), 0.9, 4 / 3, 100.0)
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 8, 6)
    q = cam.orientation.astype(np.float64)
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    th = cam.tan_half().astype(np.float64)
    for py in range(6):
        for px in range(8):
            ndc = np.array([(px + 0.5) / 8 * 2 - 1, (py + 0.5) / 6 * 2 - 1])
            v = R @ np.array([ndc[0] * th[0], ndc[1] * th[1], 1.0])
            np.testing.assert_allclose(d[py * 8 + px], v / np.linalg.norm(v), atol=1e-6)
    # camera.rs:72-86: forward faces the target, local +Y points down, right-handed.
    fwd = R @ [0, 0, 1]
    assert fwd @ (np.array([0, 0.5, 0]) - np.array([2, 1.5, -3])) / np.linalg.norm([-2, -1, 3]) > 0.9999
    assert (R @ [0, 1, 0])[1] < 0 and abs(np.linalg.det(R) - 1) < 1e-4
    np.testing.assert_array_equal(o, np.broadcast_to(cam.position, o.shape))


def test_containing_cell_matches_brute_force():
    rng = np.random.default_rng(5)
    xyz = rng.uniform(-1, 1, size=(2000, 3)).astype(np.float32)
    rad = rng.uniform(0, 0.2, 2000).astype(np.float32)
    hs = scenes.HostScene(xyz, np.ones(2000, np.float32), np.zeros((2000, 3), np.float32), 0,
                          np.zeros(2001, np.uint32), np.zeros(0, np.uint32), rad)
    for _ in range(20):
        q = rng.uniform(-1.2, 1.2, 3).astype(np.float32)
        d = ((xyz.astype(np.float64) - q) ** 2).sum(1) - rad.astype(np.float64) ** 2
        assert oracle.containing_cell(oracle.of_host(hs), q) == scenes.containing_cell(hs, q) == int(np.argmin(d))
    # ties go to the lower index (lib.rs:534-537)
    dup = scenes.HostScene(np.array([[1, 0, 0], [-1, 0, 0]], np.float32), np.ones(2, np.float32),
                           np.zeros((2, 3), np.float32), 0, np.zeros(3, np.uint32), np.zeros(0, np.uint32))
    assert oracle.containing_cell(oracle.of_host(dup), [0, 0, 0]) == 0


def test_delaunay_adjacency_is_delaunay():
    """Every tetrahedron spanned by a site and three mutually adjacent neighbours
    that Qhull reports has an empty circumsphere (checked exhaustively on 200 points)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(3)
    xyz = rng.uniform(-1, 1, size=(200, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    P = xyz.astype(np.float64)
    tri = Delaunay(P)
    for s in tri.simplices:
        A = 2 * (P[s[1:]] - P[s[0]])
        b = (P[s[1:]] ** 2).sum(1) - (P[s[0]] ** 2).sum()
        c = np.linalg.solve(A, b)
        r2 = ((P[s[0]] - c) ** 2).sum()
        d2 = ((P - c) ** 2).sum(1)
        d2[s] = np.inf
        assert d2.min() >= r2 * (1 - 1e-9)
        for a in s:
            row = nb[off[a]:off[a + 1]]
            for bb in s:
                if bb != a:
                    assert bb in row
    for i in range(200):
        row = nb[off[i]:off[i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0) and i not in row


def _compact_brute(status, cells, max_steps):
    """Per-path loops over the rows: each live step once, in path then step order."""
    slots, cs, px = [], [], []
    for p, s in enumerate(status):
        for k in range(min(int(s) & 0x7FFFFFFF, max_steps)):
            slots.append(p * max_steps + k)
            cs.append(cells[p, k])
            px.append(p)
    return np.array(slots, np.uint32), np.array(cs, np.uint32), np.array(px, np.uint32)


def test_compact_paths_matches_brute_force_and_reference_invariants():
    """path_compact.wgsl:24-50 against per-path loops, and the checks of
    gpu_path_compaction_emits_each_active_dense_record_once (tests/
    gpu_path_record.rs:1078-1196): count = total live steps, every dense slot
    emitted once, its cell, pixel = slot / max_steps, active 1."""
    rng = np.random.default_rng(5)
    L = 96
    status = rng.integers(0, L + 1, 300).astype(np.uint32)
    status[:4] = [0, L, 1, 0]
    status[10:20] |= np.uint32(oracle.TRUNCATED_BIT)  # truncated paths keep their count
    cells = rng.integers(0, 1 << 20, (300, L)).astype(np.uint32)
    r = oracle.compact_paths(status, cells, L, base=7)
    slots, cs, px = _compact_brute(status, cells, L)
    total = int((status & 0x7FFFFFFF).astype(np.int64).sum())
    assert r["count"] == 7 + total == 7 + len(slots)
    np.testing.assert_array_equal(r["dense_slots"], slots)
    np.testing.assert_array_equal(r["cells"], cs)
    np.testing.assert_array_equal(r["pixel_indices"], px)
    assert len(set(r["dense_slots"].tolist())) == total
    np.testing.assert_array_equal(r["pixel_indices"], r["dense_slots"] // L)
    np.testing.assert_array_equal(r["cells"], cells.reshape(-1)[r["dense_slots"]])
    assert (r["active"] == 1.0).all()
    e = oracle.compact_paths(np.zeros(0, np.uint32), np.zeros((0, L), np.uint32), L)
    assert e["count"] == 0 and len(e["dense_slots"]) == 0
