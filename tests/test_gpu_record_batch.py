"""bv_record_paths_batch (PathRecorder::dispatch_batch with device-generated
camera rays) against oracle.record_batch.

Bar: Voronoi rows (cells, next cells, dt, status) and the per-row ray outputs
(origin, unit direction, slice index) bit-exact; power-cell rows as the splat
recorder's (indices, dt and status bit-exact, Jacobian rows and tangent within
1e-5 relative, floored at 1).  Rows outside every slice, and slots past a
row's count, keep their previous contents.  Slices cover gaps, an empty slice,
image_pixel_offset, pixel indices past the image, a per-slice depth and dt cap
(separate launches) and more slices than one launch holds.
"""
import numpy as np
import pytest

import oracle
from bv import scenes

pytestmark = pytest.mark.gpu

SENTINEL = 0x5A5A5A5A


def _host(t, u32=False):
    a = t.cpu().numpy()
    return a.view(np.uint32) if u32 else a


def _views(cams, specs, sc):
    """specs: (pixel_offset, num_pixels, width, height, image_pixel_offset, depth, max_path_dt)."""
    import bv

    args, orc = [], []
    for c, (off, n, W, H, ipo, depth, mdt) in zip(cams, specs):
        start = oracle.containing_cell(sc, c.position) if sc.radius is None else 0
        args.append(bv.RecordPathsArgs(c, W, H, n, 0, start_point=start, pixel_offset=off, image_pixel_offset=ipo,
                                       max_path_dt=mdt, depth=depth))
        orc.append(dict(position=c.position, orientation=c.orientation, fov=c.fov, principal=c.principal,
                        image_width=W, image_height=H, image_pixel_offset=ipo, pixel_offset=off, num_pixels=n,
                        start_point=start, depth=depth, max_path_dt=mdt))
    return args, orc


def _run(cloud, args, idx, cap, L):
    import torch
    import bv

    rec = bv.PathRecorder(cloud)
    b = rec.buffers(cap, L, rays=True)
    for k in ("cells", "next_cells", "status", "view_index"):
        getattr(b, k).fill_(SENTINEL)
    b.write_pixel_indices(idx)
    for a in args:
        a.max_steps = L
    rec.dispatch_batch(b, args)
    torch.cuda.synchronize()
    return b


def _check_rows(b, ref, views, cap, L, power=False):
    covered = np.zeros(cap, bool)
    st = _host(b.status, True)
    for v in views:
        covered[v["pixel_offset"]:v["pixel_offset"] + v["num_pixels"]] = True
    np.testing.assert_array_equal(st[covered], ref["status"][covered], err_msg="status")
    assert np.all(st[~covered] == SENTINEL), "a row outside every slice was written"
    cnt = st[covered] & 0x7FFFFFFF
    live = np.zeros((cap, L), bool)
    live[covered] = np.arange(L)[None, :] < cnt[:, None]
    keys = ("cells", "next_cells") + (("previous_cells",) if power else ())
    for k in keys:
        np.testing.assert_array_equal(_host(getattr(b, k), True)[live], ref[k][live], err_msg=k)
    np.testing.assert_array_equal(_host(b.dts)[live], ref["dts"][live], err_msg="dts")
    assert np.all(_host(b.mask)[live] == 1.0)
    cells = _host(b.cells, True)
    assert np.all(cells[covered][~live[covered]] == SENTINEL), "a slot past the row count was written"
    if power:
        for k in ("dt_grad_previous", "dt_grad_current", "dt_grad_next", "dt_reference_tangent"):
            a, r = _host(getattr(b, k))[live], ref[k][live]
            err = np.abs(a - r) / np.maximum(1.0, np.abs(r))
            assert float(err.max()) <= 1e-5, f"{k}: rel err {float(err.max())}"
    np.testing.assert_array_equal(_host(b.ray_origins)[covered], ref["ray_origins"][covered])
    np.testing.assert_array_equal(_host(b.ray_dirs)[covered], ref["ray_dirs"][covered])
    vi = _host(b.view_index, True)
    np.testing.assert_array_equal(vi[covered], ref["view_index"][covered])
    assert np.all(vi[~covered] == SENTINEL)
    return int(live.sum())


@pytest.fixture(scope="module")
def uniform_cloud(c1_scene):
    import bv

    return bv.RadFoamGpuCloud(c1_scene), oracle.of_host(c1_scene)


def test_batch_voronoi_slices(uniform_cloud):
    cloud, sc = uniform_cloud
    rng = np.random.default_rng(11)
    cams = [scenes.looking_at(rng.uniform(-0.5, 0.5, 3), rng.uniform(-1, 1, 3), 0.9, 1.3, 100.0) for _ in range(4)]
    specs = [(0, 1000, 64, 48, 0, 100.0, 50.0),
             (1200, 1600, 40, 40, 7, 100.0, 50.0),   # gap 1000..1199; 7 + index can pass the image
             (2800, 0, 16, 16, 0, 100.0, 50.0),      # empty slice
             (2800, 500, 32, 24, 0, 0.8, 0.02)]      # own depth and dt cap: a second launch
    args, views = _views(cams, specs, sc)
    cap, L = 3400, 64
    idx = np.zeros(cap, np.uint32)
    idx[0:1000] = rng.permutation(64 * 48)[:1000]
    idx[1200:2800] = rng.integers(0, 40 * 40, 1600)
    idx[2800:3300] = rng.integers(0, 32 * 24, 500)
    b = _run(cloud, args, idx, cap, L)
    ref = oracle.record_batch(sc, views, idx, capacity=cap, max_steps=L)
    n = _check_rows(b, ref, views, cap, L)
    print(f"voronoi batch: {n} live steps")
    st = _host(b.status, True)
    assert np.all(st[2800:3300] & 0x7FFFFFFF > 0)
    assert np.all(_host(b.dts)[2800:3300][_host(b.mask)[2800:3300] == 1.0] <= np.float32(0.02))


def test_batch_more_slices_than_one_launch(uniform_cloud):
    cloud, sc = uniform_cloud
    rng = np.random.default_rng(12)
    cams = [scenes.looking_at(rng.uniform(-0.5, 0.5, 3), rng.uniform(-1, 1, 3), 0.7, 1.0, 100.0) for _ in range(31)]
    specs = [(70 * k, 64 + (k % 5), 12, 12, k, 100.0, 50.0) for k in range(31)]
    args, views = _views(cams, specs, sc)
    cap, L = 70 * 31, 48
    idx = rng.integers(0, 144, cap).astype(np.uint32)
    b = _run(cloud, args, idx, cap, L)
    ref = oracle.record_batch(sc, views, idx, capacity=cap, max_steps=L)
    _check_rows(b, ref, views, cap, L)


@pytest.mark.parametrize("lanes", [None, "4"], ids=["default", "g4"])
def test_batch_large_slices(uniform_cloud, lanes, monkeypatch):
    """Six slices of 32,768 random pixels each (196,608 rows): rows, status
    words and ray outputs bit-exact, slots past each row's count untouched."""
    print(f"case large batch recorder lanes={lanes}", flush=True)
    if lanes is not None:
        monkeypatch.setenv("BV_LANES_PER_RAY", lanes)
    cloud, sc = uniform_cloud
    rng = np.random.default_rng(14)
    cams = [scenes.looking_at(rng.uniform(-0.5, 0.5, 3), rng.uniform(-1, 1, 3), 0.9, 1.3, 100.0) for _ in range(6)]
    per = 32_768
    specs = [(per * k, per, 256, 192, 0, 100.0, 50.0) for k in range(6)]
    args, views = _views(cams, specs, sc)
    cap, L = per * 6, 48
    idx = rng.integers(0, 256 * 192, cap).astype(np.uint32)
    b = _run(cloud, args, idx, cap, L)
    ref = oracle.record_batch(sc, views, idx, capacity=cap, max_steps=L)
    n = _check_rows(b, ref, views, cap, L)
    assert n > cap


def test_batch_power_cells():
    import bv

    host = scenes.powerfoam_scene(4096, seed=6, sh_degree=0)
    cloud = bv.RadFoamGpuCloud(host)
    sc = oracle.of_host(host)
    cams = scenes.ring_cameras(3, 2.6, 0.8, 1.25, 20.0)
    specs = [(0, 600, 40, 32, 0, 20.0, 50.0), (700, 500, 24, 24, 3, 20.0, 50.0), (1200, 400, 30, 20, 0, 20.0, 0.01)]
    args, views = _views(cams, specs, sc)
    rng = np.random.default_rng(13)
    cap, L = 1650, 40
    idx = np.zeros(cap, np.uint32)
    idx[0:600] = rng.permutation(40 * 32)[:600]
    idx[700:1200] = rng.integers(0, 24 * 24, 500)
    idx[1200:1600] = rng.integers(0, 30 * 20, 400)
    b = _run(cloud, args, idx, cap, L)
    ref = oracle.record_batch(sc, views, idx, capacity=cap, max_steps=L)
    n = _check_rows(b, ref, views, cap, L, power=True)
    assert n > 0
    print(f"power batch: {n} live steps")


def test_batch_guards(uniform_cloud):
    import bv
    from bv._lib import BvError

    cloud, sc = uniform_cloud
    cam = scenes.looking_at((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.9, 1.0, 100.0)
    rec = bv.PathRecorder(cloud)
    b = rec.buffers(100, 8)
    ok = dict(image_width=10, image_height=10, num_pixels=40, max_steps=8, start_point=0)
    bad = [
        [bv.RecordPathsArgs(cam, **ok), bv.RecordPathsArgs(cam, **dict(ok, pixel_offset=30))],   # overlap
        [bv.RecordPathsArgs(cam, **dict(ok, pixel_offset=50)), bv.RecordPathsArgs(cam, **ok)],   # out of order
        [bv.RecordPathsArgs(cam, **dict(ok, pixel_offset=70))],                                  # past capacity
        [bv.RecordPathsArgs(cam, **dict(ok, image_width=0))],                                    # empty image
        [bv.RecordPathsArgs(cam, **dict(ok, start_point=cloud.n_sites))],                        # bad start
    ]
    for args in bad:
        with pytest.raises(BvError):
            rec.dispatch_batch(b, args)
    with pytest.raises(ValueError):
        rec.dispatch_batch(b, [bv.RecordPathsArgs(cam, **dict(ok, max_steps=9))])
    rec.dispatch_batch(b, [])
