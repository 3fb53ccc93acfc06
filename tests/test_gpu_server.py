"""The serving loop on the GPU: each image is bit-identical alone and in a mix of
concurrent requests (and to the direct render with the same lanes per ray),
admission waits for memory, oversized requests are refused, stats are sane."""
import numpy as np
import pytest

from bv import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def clouds():
    import bv

    pw = scenes.power_scene(20_000, seed=6)
    vo = scenes.uniform_scene(8_000, seed=3)
    return {"power": (pw, bv.RadFoamGpuCloud(pw)), "voronoi": (vo, bv.RadFoamGpuCloud(vo))}


def _cam(k):
    return scenes.looking_at((3.0 * np.cos(k), 0.3, 3.0 * np.sin(k)), (0, 0, 0), np.deg2rad(55), 4 / 3, 100.0)


def test_alone_equals_mixed_and_direct(clouds, monkeypatch):
    import bv

    print("case server alone vs mixed", flush=True)
    rng = np.random.default_rng(0)
    lights = scenes.random_lights(3, rng)
    srv = bv.Server(hbm_budget_bytes=64 << 20, max_active=16)
    ids = {k: srv.add_scene(c) for k, (_, c) in clouds.items()}
    probe = dict(scene=ids["power"], camera=_cam(0.3), size=(96, 72), mode=bv.EVAL_RELIGHT, lights=lights,
                 visibility=True)
    probe_sh = dict(scene=ids["voronoi"], camera=_cam(1.1), size=(80, 64), mode=bv.EVAL_SH)
    t, alone = srv.submit(**probe)
    srv.wait(t)
    t, alone_sh = srv.submit(**probe_sh)
    srv.wait(t)
    # the same requests among a stream of others of mixed size, scene and mode
    tickets = []
    for k in range(12):
        if k % 3 == 0:
            tickets.append(srv.submit(ids["voronoi"], _cam(k), (int(rng.integers(16, 200)), 50), bv.EVAL_SH)[0])
        else:
            tickets.append(srv.submit(ids["power"], _cam(k), (64, int(rng.integers(16, 160))), bv.EVAL_RELIGHT,
                                      scenes.random_lights(int(rng.integers(1, 9)), rng), bool(k % 2))[0])
        if k == 5:
            t1, mixed = srv.submit(**probe)
            t2, mixed_sh = srv.submit(**probe_sh)
    for t in tickets + [t1, t2]:
        srv.wait(t)
    np.testing.assert_array_equal(alone, mixed)
    np.testing.assert_array_equal(alone_sh, mixed_sh)
    assert alone[:, 3].max() > 0.5 and alone[:, :3].max() > 0
    # ... and equal to the direct renders at the serving kernel's 4 lanes per ray
    monkeypatch.setenv("BV_LANES_PER_RAY", "4")
    direct = bv.relight_camera(clouds["power"][1], probe["camera"], probe["size"], lights, visibility=True)
    np.testing.assert_array_equal(direct.cpu().numpy(), alone)
    tr = bv.RadFoamGpuTracer(clouds["voronoi"][1]).dispatch(probe_sh["camera"], probe_sh["size"])
    np.testing.assert_array_equal(tr["rgba"].cpu().numpy(), alone_sh)
    st = srv.stats()
    assert st["completed"] == 16 and st["p50_ms"] <= st["p99_ms"] and st["mrays_per_s"] > 0
    srv.close()


def test_admission_by_memory_and_capacity(clouds):
    import bv

    print("case server admission", flush=True)
    pw, cloud = clouds["power"]
    # budget for one 256x256 image (1 MiB) at a time
    srv = bv.Server(hbm_budget_bytes=1 << 20, max_active=8)
    sid = srv.add_scene(cloud)
    lights = scenes.random_lights(2, np.random.default_rng(1))
    outs = [srv.submit(sid, _cam(k), (256, 256), bv.EVAL_RELIGHT, lights) for k in range(4)]
    for t, _ in outs:
        srv.wait(t)
    st = srv.stats()
    assert st["completed"] == 4 and st["peak_bytes"] <= 1 << 20
    # a completed wait released its ticket
    with pytest.raises(bv.BvError) as e:
        srv.poll(outs[0][0])
    assert e.value.code == 1  # BV_ERR_INVALID
    with pytest.raises(bv.BvError) as e:
        srv.submit(sid, _cam(0), (512, 512), bv.EVAL_RELIGHT, lights)
    assert e.value.code == 5  # BV_ERR_CAPACITY
    ref = bv.Server(hbm_budget_bytes=8 << 20, max_active=8)
    rid = ref.add_scene(cloud)
    for k, (t, out) in enumerate(outs):
        t2, again = ref.submit(rid, _cam(k), (256, 256), bv.EVAL_RELIGHT, lights)
        ref.wait(t2)
        np.testing.assert_array_equal(out, again)
    ref.close()
    srv.close()


def test_per_light_visibility_served(clouds, monkeypatch):
    """Per-light visibility rays through the serving loop equal the direct
    relit render with the same per-light flags (4 lanes per ray)."""
    import bv

    print("case server per-light visibility", flush=True)
    rng = np.random.default_rng(9)
    lights = scenes.random_lights(4, rng)
    lights[1]["visibility"] = True
    lights[2]["visibility"] = True
    srv = bv.Server(hbm_budget_bytes=16 << 20, max_active=4)
    sid = srv.add_scene(clouds["power"][1])
    t, img = srv.submit(sid, _cam(0.9), (80, 60), bv.EVAL_RELIGHT, lights)
    srv.wait(t)
    srv.close()
    monkeypatch.setenv("BV_LANES_PER_RAY", "4")
    direct = bv.relight_camera(clouds["power"][1], _cam(0.9), (80, 60), lights).cpu().numpy()
    np.testing.assert_array_equal(direct, img)
    none = bv.relight_camera(clouds["power"][1], _cam(0.9), (80, 60),
                             [dict(light, visibility=False) for light in lights]).cpu().numpy()
    assert not np.array_equal(none, img)
