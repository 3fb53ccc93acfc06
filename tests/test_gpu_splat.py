"""HIP PowerFoam splat tracer against the oracle (trace_powerfoam_splats,
trace.rs:1331-1546).

Bar: recorded rows (cell, near, dt, count) are bit-exact; for traces the
composited cells (steps, last cell, peak cell) are bit-exact and colour,
transmittance, t_end, depth and peak weight agree within 1e-5 absolute in FP32
(the oracle uses libm expf, the device its own exp).
"""
import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5
ALL_OUT = ("rgba", "steps", "last_point", "t_end", "depth_mode", "peak_weight", "peak_point")


def _bv():
    import bv

    return bv


def dev(a, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t, view=None):
    a = t.cpu().numpy()
    return a.view(view) if view is not None else a


def compare_trace(got, ref, label, rows=None):
    sel = slice(None) if rows is None else rows
    np.testing.assert_array_equal(host(got["steps"], np.uint32)[sel], ref["steps"][sel], err_msg=f"{label}: steps")
    np.testing.assert_array_equal(host(got["last_point"], np.uint32)[sel], ref["last_point"][sel],
                                  err_msg=f"{label}: last")
    np.testing.assert_array_equal(host(got["peak_point"], np.uint32)[sel], ref["peak_point"][sel],
                                  err_msg=f"{label}: peak cell")
    for k in ("rgba", "t_end", "depth_mode", "peak_weight"):
        err = float(np.abs(host(got[k])[sel] - ref[k][sel]).max()) if ref[k].size else 0.0
        assert err <= TOL, f"{label}: {k} differs by {err}"


def compare_record(got, ref, label):
    cnt = host(got["count"], np.uint32)
    np.testing.assert_array_equal(cnt, ref["count"], err_msg=f"{label}: count")
    cells, nears, dts = host(got["cells"], np.uint32), host(got["nears"]), host(got["dts"])
    for k in range(len(cnt)):
        c = int(cnt[k])
        np.testing.assert_array_equal(cells[k, :c], ref["cells"][k, :c], err_msg=f"{label}: cells row {k}")
        np.testing.assert_array_equal(nears[k, :c], ref["nears"][k, :c], err_msg=f"{label}: near row {k}")
        np.testing.assert_array_equal(dts[k, :c], ref["dts"][k, :c], err_msg=f"{label}: dt row {k}")


def splat_both(scene, cloud, o, d, **kw):
    bv = _bv()
    st = bv.TraceSettings(weight_threshold=kw.get("weight_threshold", 0.001), max_steps=kw.get("max_steps", 1024),
                          depth=kw.get("depth", 100.0), eval_mode=kw.get("eval_mode", bv.EVAL_SH),
                          const_rgb=kw.get("const_rgb", (1.0, 1.0, 1.0)))
    got = bv.splat_trace_rays(cloud, dev(o), dev(d), settings=st, outputs=ALL_OUT)
    ref = oracle.trace_splats(oracle.of_host(scene), o, d, 0, weight_threshold=st.weight_threshold,
                              max_steps=st.max_steps, depth=st.depth, eval_mode=st.eval_mode,
                              const_rgb=st.const_rgb)
    return got, ref


def _fixture(xyz, radius, density, sh, deg=0):
    n = len(xyz)
    xyz = np.asarray(xyz, np.float32)
    radius = np.asarray(radius, np.float32)
    off, nb = scenes.cech_csr(xyz, radius)
    return scenes.HostScene(xyz, np.asarray(density, np.float32), np.asarray(sh, np.float32).reshape(n, -1), deg, off,
                            nb, radius)


@pytest.fixture(scope="module")
def pf_scene():
    """6,000 PowerFoam sites (cube of side 0.9, radii log-uniform [1e-3, 1e-1],
    Cech adjacency), SH degree 2, density 3 exp(N(0,1))."""
    bv = _bv()
    h = scenes.powerfoam_scene(6000, seed=7, sh_degree=2, density_scale=3.0)
    return h, bv.RadFoamGpuCloud(h)


def _rays_at(scene_half, m, rng, origin=(0.0, 0.1, 1.6)):
    o = np.tile(np.asarray(origin, np.float32), (m, 1))
    tgt = rng.uniform(-scene_half, scene_half, (m, 3))
    d = (tgt - o).astype(np.float32)
    return o, d


def test_disconnected_supports():
    bv = _bv()
    print("case splat two supports", flush=True)
    h = scenes.HostScene(np.array([[0, 0, 3], [0, 0, 6]], np.float32), np.ones(2, np.float32),
                         np.zeros((2, 3), np.float32), 0, np.zeros(3, np.uint32), np.zeros(0, np.uint32),
                         np.array([0.5, 0.5], np.float32))
    cloud = bv.RadFoamGpuCloud(h)
    o = np.zeros((1, 3), np.float32)
    d = np.array([[0, 0, 1]], np.float32)
    rec = bv.splat_record(cloud, dev(o), dev(d), 8, depth=10.0)
    assert host(rec["count"])[0] == 2
    assert list(host(rec["cells"])[0, :2]) == [0, 1]
    np.testing.assert_allclose(host(rec["dts"])[0, :2], 1.0, atol=1e-6)
    got, ref = splat_both(h, cloud, o, d, depth=10.0)
    compare_trace(got, ref, "two supports")


def test_single_support_camera_rows():
    """The batched-rows fixture (tests/radfoam_gpu_vs_cpu.rs:665-779) with the
    radius of the single-support test beside it (:580-601, radius 1.2; at the
    fixture's radius 4 the camera lies inside the 4 r exclusion and the image
    is empty): one site at z = 5, a 9x3 camera.  The image equals the oracle on
    the same camera rays and equals a ray-batch trace of those rays."""
    bv = _bv()
    print("case splat single support camera", flush=True)
    h = scenes.HostScene(np.array([[0, 0, 5]], np.float32), np.array([1.0], np.float32),
                         np.array([[0.4, -0.2, 0.1]], np.float32), 0, np.zeros(2, np.uint32), np.zeros(0, np.uint32),
                         np.array([1.2], np.float32))
    cloud = bv.RadFoamGpuCloud(h)
    cam = scenes.Camera(np.zeros(3, np.float32), 10.0, np.array([0, 0, 0, 1], np.float32),
                        np.array([0.5, 0.5], np.float32))
    tracer = bv.PowerFoamGpuSplatTracer(cloud, bv.RadFoamTraceSettings(max_steps=4, weight_threshold=1e-4), (9, 3))
    img = tracer.dispatch(cam, outputs=ALL_OUT)
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 9, 3)
    ref = oracle.trace_splats(oracle.of_host(h), o, d, 0, weight_threshold=1e-4, max_steps=4, depth=10.0)
    compare_trace(img, ref, "9x3 camera")
    assert host(img["rgba"])[:, 3].max() > 0.5
    rays = bv.splat_trace_rays(cloud, dev(o), dev(d), settings=bv.TraceSettings(1e-4, 4, 10.0, bv.EVAL_SH),
                               outputs=ALL_OUT)
    for k in ALL_OUT:
        np.testing.assert_array_equal(host(rays[k]), host(img[k]), err_msg=k)


def test_record_parity(pf_scene):
    bv = _bv()
    h, cloud = pf_scene
    print("case splat record 6000", flush=True)
    o, d = _rays_at(0.45, 512, np.random.default_rng(1))
    got = bv.splat_record(cloud, dev(o), dev(d), 256, depth=100.0)
    ref = oracle.record_splats(oracle.of_host(h), o, d, max_steps=256, depth=100.0)
    assert ref["count"].max() > 16  # rows longer than one k-buffer round
    compare_record(got, ref, "record L=256")
    got = bv.splat_record(cloud, dev(o), dev(d), 5, depth=100.0)
    ref = oracle.record_splats(oracle.of_host(h), o, d, max_steps=5, depth=100.0)
    compare_record(got, ref, "record L=5")


def test_record_parity_long_rows():
    """60,000 sites in [-1,1]^3, radii log-uniform [0.01, 0.1]: rows of 40+
    intervals, so three or more k-buffer rounds per ray."""
    bv = _bv()
    print("case splat record 60000 long rows", flush=True)
    rng = np.random.default_rng(7)
    n = 60000
    xyz = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(0.01), np.log(0.1), n)).astype(np.float32)
    off, nb = scenes.cech_csr(xyz, r)
    h = scenes.HostScene(xyz, np.ones(n, np.float32), np.zeros((n, 3), np.float32), 0, off, nb, r)
    cloud = bv.RadFoamGpuCloud(h)
    o, d = _rays_at(0.9, 256, rng, origin=(0.0, 0.1, 2.5))
    got = bv.splat_record(cloud, dev(o), dev(d), 512, depth=100.0)
    ref = oracle.record_splats(oracle.of_host(h), o, d, max_steps=512, depth=100.0)
    assert ref["count"].max() > 40
    compare_record(got, ref, "record 60000")


@pytest.mark.parametrize("mode", ["opacity", "constant", "sh"])
@pytest.mark.parametrize("limits", [(0.001, 1024), (0.0, 1024), (0.001, 3), (0.3, 1024)],
                         ids=["default", "no-early-out", "max3", "thr0.3"])
def test_trace_parity(pf_scene, mode, limits):
    bv = _bv()
    h, cloud = pf_scene
    thr, ms = limits
    print(f"case splat trace {mode} thr={thr} max={ms}", flush=True)
    em = {"opacity": bv.EVAL_OPACITY, "constant": bv.EVAL_CONSTANT, "sh": bv.EVAL_SH}[mode]
    o, d = _rays_at(0.45, 1024, np.random.default_rng(2))
    got, ref = splat_both(h, cloud, o, d, eval_mode=em, weight_threshold=thr, max_steps=ms,
                          const_rgb=(0.2, 0.5, 0.9))
    compare_trace(got, ref, f"{mode} thr={thr} max={ms}")
    assert ref["rgba"][:, 3].max() > 0.5


def test_camera_parity(pf_scene):
    bv = _bv()
    h, cloud = pf_scene
    print("case splat camera 64x48", flush=True)
    cam = scenes.looking_at((1.2, 0.4, 1.1), (0, 0, 0), np.deg2rad(50), 4 / 3, 100.0)
    tracer = bv.PowerFoamGpuSplatTracer(cloud)
    img = tracer.dispatch(cam, (64, 48), outputs=ALL_OUT)
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 64, 48)
    ref = oracle.trace_splats(oracle.of_host(h), o, d, 0, depth=100.0)
    compare_trace(img, ref, "camera 64x48")
    assert ref["rgba"][:, 3].mean() > 0.2


def test_edge_rays(pf_scene):
    bv = _bv()
    h, cloud = pf_scene
    print("case splat edge rays", flush=True)
    o = np.array([[0, 0, 1.6], [0, 0, 1.6], [5, 5, 5], [0.0, 0.0, 0.0], [0, 0, 1.6]], np.float32)
    d = np.array([[0, 0, 0], [0, 0, 1], [1, 1, 1], [0.3, -0.2, 1.0], [1e-30, 0, -1]], np.float32)
    got, ref = splat_both(h, cloud, o, d)
    compare_trace(got, ref, "edge rays")
    assert host(got["steps"])[0] == 0 and host(got["rgba"])[0].max() == 0
    # an empty ray batch is a no-op
    e = np.zeros((0, 3), np.float32)
    out = bv.splat_trace_rays(cloud, dev(e), dev(e))
    assert out["rgba"].shape == (0, 4)


def test_scene_update_rebuilds_bvh(pf_scene):
    import torch

    bv = _bv()
    h0, _ = pf_scene
    print("case splat update", flush=True)
    h = scenes.HostScene(h0.xyz.copy(), h0.density.copy(), h0.sh.copy(), h0.sh_degree, h0.offsets, h0.neighbors,
                         h0.radius)
    cloud = bv.RadFoamGpuCloud(h)
    o, d = _rays_at(0.45, 256, np.random.default_rng(4))
    bv.splat_record(cloud, dev(o), dev(d), 64)  # builds the BVH on the old geometry
    rng = np.random.default_rng(9)
    moved = (h.xyz + rng.normal(scale=0.05, size=h.xyz.shape)).astype(np.float32)
    cloud.update(xyz=dev(moved))
    torch.cuda.synchronize()
    h2 = scenes.HostScene(moved, h.density, h.sh, h.sh_degree, h.offsets, h.neighbors, h.radius)
    got = bv.splat_record(cloud, dev(o), dev(d), 64)
    ref = oracle.record_splats(oracle.of_host(h2), o, d, max_steps=64, depth=10_000.0)
    compare_record(got, ref, "after update")


def test_walk_scene_rejected():
    bv = _bv()
    h = scenes.uniform_scene(512, seed=1)
    cloud = bv.RadFoamGpuCloud(h)
    o = np.zeros((1, 3), np.float32)
    with pytest.raises(bv.BvError) as e:
        bv.splat_trace_rays(cloud, dev(o), dev(o + 1))
    assert e.value.code == 2


def test_larger_scene_parity():
    bv = _bv()
    print("case splat 65536", flush=True)
    h = scenes.powerfoam_scene(65536, seed=4, sh_degree=0)
    cloud = bv.RadFoamGpuCloud(h)
    cam = scenes.looking_at((2.2, 0.8, 2.0), (0, 0, 0), np.deg2rad(45), 1.0, 100.0)
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 24, 24)
    got = bv.splat_record(cloud, dev(o), dev(d), 128, depth=100.0)
    ref = oracle.record_splats(oracle.of_host(h), o, d, max_steps=128, depth=100.0)
    compare_record(got, ref, "65536 record")
    got, ref = splat_both(h, cloud, o, d, depth=100.0)
    compare_trace(got, ref, "65536 trace")
    assert ref["rgba"][:, 3].mean() > 0.5


def test_record_at_1m_sites():
    """The recorder at 1,048,576 sites (Čech adjacency from bv_compute_cech)
    over one 1024x768 camera's rays, 64-interval rows, depth 10,000: 256 rays
    spread over the image against the oracle bit for bit, and over all 786,432
    rows the properties that hold at any size: counts within L, rows in
    increasing (near, cell) order with positive dt, and the Jacobian recorder
    writing the same cells and lengths (its status adds the truncated bit when
    a 65th interval follows)."""
    import math

    import torch

    bv = _bv()
    print("case splat record 1M sites, 1024x768 camera", flush=True)
    n, W, H, L = 1 << 20, 1024, 768, 64
    h = scenes.powerfoam_scene(n, seed=6, cech=bv.compute_cech)
    cloud = bv.RadFoamGpuCloud(h)
    half = (n / 65536.0) ** (1.0 / 3.0)
    cam = scenes.looking_at((2.4 * half, 0.3 * half, 0.5 * half), (0, 0, 0), math.radians(55.0), 4 / 3, 100.0)
    o, d = bv.camera_rays(cam, W, H)
    got = bv.splat_record(cloud, o, d, L)
    jac = bv.splat_record_jacobians(cloud, o, d, L)
    torch.cuda.synchronize()

    pick = np.linspace(0, W * H - 1, 256).astype(np.int64)
    oh, dh = host(o)[pick], host(d)[pick]
    ref = oracle.record_splats(oracle.of_host(h), oh, dh, max_steps=L)
    sub = {k: got[k][torch.from_numpy(pick).cuda()] for k in ("cells", "nears", "dts", "count")}
    compare_record(sub, ref, "1M sites subsample")
    assert ref["count"].max() == L and ref["count"].min() < L  # full and partial rows both checked

    cnt = got["count"].to(torch.int64)
    assert int(cnt.max()) <= L and int(cnt.min()) >= 0
    live = torch.arange(L, device=cnt.device)[None, :] < cnt[:, None]
    nears, dts = got["nears"], got["dts"]
    cells = got["cells"].to(torch.int64) & 0xFFFFFFFF
    assert bool(torch.all(dts[live] > 0))
    assert bool(torch.all(cells[live] < n))
    # (near, cell) strictly increasing along each row's live slots
    pair = live[:, 1:]
    n0, n1 = nears[:, :-1], nears[:, 1:]
    c0, c1 = cells[:, :-1], cells[:, 1:]
    ordered = (n1 > n0) | ((n1 == n0) & (c1 > c0))
    assert bool(torch.all(ordered[pair]))
    status = jac["status"].to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(status & 0x7FFFFFFF, cnt)
    truncated = (status >> 31) == 1
    assert bool(torch.all(cnt[truncated] == L))
    jc = jac["cells"].to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(jc[live], cells[live])
    cloud.deinit()


@pytest.fixture(scope="module")
def pf_lit():
    """8,000 PowerFoam sites with materials (SH degree 0), density 20 exp(N(0,1))."""
    bv = _bv()
    h = scenes.powerfoam_scene(8000, seed=11, sh_degree=0)
    return h, bv.RadFoamGpuCloud(h)


@pytest.mark.parametrize("nlights,vis", [(1, False), (1, True), (3, True), (8, True), (5, "per-light")])
def test_splat_relight_matches_oracle(pf_lit, nlights, vis):
    bv = _bv()
    h, cloud = pf_lit
    print(f"case splat relight lights={nlights} vis={vis}", flush=True)
    rng = np.random.default_rng(20 + nlights)
    L = scenes.random_lights(nlights, rng)
    if vis == "per-light":
        # only the even lights cast visibility rays
        for i, light in enumerate(L):
            light["visibility"] = i % 2 == 0
        vis = False
    cam = scenes.looking_at((0.4, 0.3, 2.2), (0, 0, 0), np.deg2rad(50), 1.0, 100.0)
    W = H = 40
    got = bv.splat_relight_camera(cloud, cam, (W, H), L, visibility=vis).cpu().numpy()
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, W, H)
    ref = oracle.trace_splat_relight(oracle.of_host(h), h.material, o, d, L, visibility=vis, depth=100.0)
    # Shadow traces stop once transmittance <= 0.001; device and host exp may
    # place that stop one interval apart, which moves a light's visibility by at
    # most 0.001: the bound adds 0.001 x the pixel's unshadowed light sum.
    bound = 2e-5 * np.maximum(1.0, np.abs(ref))
    if vis or any(light.get("visibility") for light in L):
        unlit = [dict(light, visibility=False) for light in L]
        free = oracle.trace_splat_relight(oracle.of_host(h), h.material, o, d, unlit, visibility=False, depth=100.0)
        bound = bound + 1.001e-3 * np.abs(free)
    excess = np.abs(got - ref) - bound
    assert excess.max() <= 0, float(np.abs(got - ref).max())
    close = np.abs(got - ref) <= 2e-5 * np.maximum(1.0, np.abs(ref))
    assert close.mean() >= 0.99
    assert ref[:, 3].max() > 0.5 and ref[:, :3].max() > 0
    # rays entry point equals the camera entry point
    rays = bv.splat_relight_rays(cloud, dev(o), dev(d), L, visibility=vis,
                                 settings=bv.TraceSettings(depth=100.0)).cpu().numpy()
    np.testing.assert_array_equal(rays, got)


def test_splat_serving_alone_equals_mixed(pf_lit, pf_scene):
    bv = _bv()
    print("case splat serving", flush=True)
    h, cloud = pf_lit
    _, cloud_sh = pf_scene
    srv = bv.Server(hbm_budget_bytes=32 << 20, max_active=16)
    a = srv.add_scene(cloud)
    b = srv.add_scene(cloud_sh)
    rng = np.random.default_rng(3)
    L = scenes.random_lights(4, rng)
    cam = lambda k: scenes.looking_at((1.6 * np.cos(k), 0.4, 1.6 * np.sin(k)), (0, 0, 0), np.deg2rad(55), 4 / 3,
                                      100.0)
    probe = dict(scene=a, camera=cam(0.7), size=(72, 56), mode=bv.EVAL_RELIGHT, lights=L, visibility=True,
                 renderer=bv.RENDER_SPLAT)
    probe_sh = dict(scene=b, camera=cam(2.0), size=(64, 40), mode=bv.EVAL_SH, renderer=bv.RENDER_SPLAT)
    t, alone = srv.submit(**probe)
    srv.wait(t)
    t, alone_sh = srv.submit(**probe_sh)
    srv.wait(t)
    tickets = []
    for k in range(10):
        if k % 3 == 0:
            tickets.append(srv.submit(b, cam(k), (int(rng.integers(16, 120)), 48), bv.EVAL_SH,
                                      renderer=bv.RENDER_SPLAT)[0])
        elif k % 3 == 1:
            tickets.append(srv.submit(a, cam(k), (48, int(rng.integers(16, 100))), bv.EVAL_RELIGHT,
                                      scenes.random_lights(int(rng.integers(1, 9)), rng), bool(k % 2),
                                      renderer=bv.RENDER_SPLAT)[0])
        else:  # a walk request of the same scene in the same mix
            tickets.append(srv.submit(a, cam(k), (40, 40), bv.EVAL_RELIGHT, L)[0])
        if k == 4:
            t1, mixed = srv.submit(**probe)
            t2, mixed_sh = srv.submit(**probe_sh)
    for t in tickets + [t1, t2]:
        srv.wait(t)
    np.testing.assert_array_equal(alone, mixed)
    np.testing.assert_array_equal(alone_sh, mixed_sh)
    direct = bv.splat_relight_camera(cloud, probe["camera"], probe["size"], L, visibility=True).cpu().numpy()
    np.testing.assert_array_equal(direct, alone)
    d_sh = bv.PowerFoamGpuSplatTracer(cloud_sh).dispatch(probe_sh["camera"], probe_sh["size"])["rgba"]
    np.testing.assert_array_equal(d_sh.cpu().numpy(), alone_sh)
    assert alone[:, 3].max() > 0.5 and alone[:, :3].max() > 0
    st = srv.stats()
    assert st["completed"] == 14
    srv.close()


def test_splat_serving_mixed_tree_depths(pf_lit):
    """Splat requests for two scenes of different sphere-tree depth served
    together (one launch's LDS stack covers the deeper tree): both images
    equal the direct splat renders."""
    bv = _bv()
    print("case splat serving mixed tree depths", flush=True)
    h, cloud = pf_lit
    big = scenes.powerfoam_scene(60_000, seed=12, sh_degree=0, adjacency=False)
    cloud_big = bv.RadFoamGpuCloud(big)
    srv = bv.Server(hbm_budget_bytes=64 << 20, max_active=8)
    a = srv.add_scene(cloud)
    b = srv.add_scene(cloud_big)
    L = scenes.random_lights(3, np.random.default_rng(4))
    cam = scenes.looking_at((1.7, 0.3, 0.9), (0, 0, 0), np.deg2rad(55), 4 / 3, 100.0)
    ta, ia = srv.submit(a, cam, (512, 384), bv.EVAL_RELIGHT, L, True, renderer=bv.RENDER_SPLAT)
    tb, ib = srv.submit(b, cam, (96, 72), bv.EVAL_RELIGHT, L, True, renderer=bv.RENDER_SPLAT)
    srv.wait(ta)
    srv.wait(tb)
    st = srv.stats()
    srv.close()
    assert st["completed"] == 2
    da = bv.splat_relight_camera(cloud, cam, (512, 384), L, visibility=True).cpu().numpy()
    db = bv.splat_relight_camera(cloud_big, cam, (96, 72), L, visibility=True).cpu().numpy()
    np.testing.assert_array_equal(da, ia)
    np.testing.assert_array_equal(db, ib)
    cloud_big.deinit()
