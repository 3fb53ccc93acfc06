"""Depth tracers (gpu/radfoam_depth.rs, gpu/powerfoam_depth.rs) against the oracle.

Both write the reference's rgba32float depth texture, per pixel (mode depth,
accumulated alpha, peak segment weight, 1): radfoam_depth.wgsl:29-35 over the
camera-seeded walk with record_depth (radfoam_trace.wgsl:164-191, density only,
depth = the middle of the heaviest segment), powerfoam_depth.wgsl:118-147 over
the independently clipped supports (depth = the heaviest segment's entry depth
+ dt / 2).

Bar: within 1e-5 absolute in FP32 of the oracle's opacity walk on the same
camera rays (libm expf against the device exp), the fourth channel exactly 1,
and the map unchanged when the SH coefficients change (the depth walk never
evaluates colour).
"""
import numpy as np
import pytest

import oracle
from bv import scenes
from conftest import c1_camera
import fixtures

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _bv():
    import bv

    return bv


def _check(img, ref, label):
    img = img.cpu().numpy()
    h, w, _ = img.shape
    flat = img.reshape(h * w, 4)
    for k, (name, want) in enumerate([("depth_mode", ref["depth_mode"]), ("alpha", ref["rgba"][:, 3]),
                                      ("peak_weight", ref["peak_weight"])]):
        err = float(np.abs(flat[:, k] - want).max())
        assert err <= TOL, f"{label}: {name} differs by {err}"
    np.testing.assert_array_equal(flat[:, 3], np.ones(h * w, np.float32))


@pytest.mark.parametrize("kernel", [0, 1], ids=["plain", "fast"])
def test_radfoam_depth_matches_oracle(c1_scene, kernel):
    """16,384 uniform sites, 96x80 camera at z = 3, start at the camera's containing cell."""
    print(f"case radfoam depth c1 kernel={kernel}", flush=True)
    bv = _bv()
    cloud = bv.RadFoamGpuCloud(c1_scene)
    cam = c1_camera()
    start = cloud.containing_point(cam.position)
    img = bv.RadFoamGpuDepthTracer(cloud, bv.RadFoamDepthSettings(), kernel).dispatch(cam, (96, 80))
    assert tuple(img.shape) == (80, 96, 4)
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 96, 80)
    ref = oracle.trace_rays(oracle.of_host(c1_scene), o, d, start, weight_threshold=0.001, max_steps=1024,
                            depth=100.0, eval_mode=bv.EVAL_OPACITY)
    _check(img, ref, "c1 96x80")
    assert ref["rgba"][:, 3].mean() > 0.2 and ref["peak_weight"].max() > 0.0
    cloud.deinit()


def test_radfoam_depth_settings_and_colour_independence():
    """A short step budget and a high weight threshold reach the walk; new SH
    coefficients leave the map bit-identical; an empty map is refused."""
    print("case radfoam depth settings", flush=True)
    bv = _bv()
    sc = fixtures.branching(sh_degree=3)
    cam = scenes.Camera(np.array([0.0, 0.0, -1.0], np.float32), 100.0, np.array([0, 0, 0, 1], np.float32),
                        np.array([1.0, 1.0], np.float32), np.array([0.08, -0.04], np.float32))
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 24, 20)
    cloud = bv.RadFoamGpuCloud(sc)
    for max_steps, wt in [(1024, 1e-4), (3, 1e-4), (1024, 0.5)]:
        tracer = bv.RadFoamGpuDepthTracer(cloud, bv.RadFoamDepthSettings(max_steps=max_steps, weight_threshold=wt))
        img = tracer.dispatch(cam, (24, 20), start_cell=0)
        ref = oracle.trace_rays(oracle.of_host(sc), o, d, 0, weight_threshold=wt, max_steps=max_steps, depth=100.0,
                                eval_mode=bv.EVAL_OPACITY)
        _check(img, ref, f"branching max_steps={max_steps} threshold={wt}")
    base = bv.RadFoamGpuDepthTracer(cloud).dispatch(cam, (24, 20), start_cell=0).cpu().numpy()
    assert base[..., 1].max() > 0.05
    sc2 = fixtures.branching(sh_degree=3)
    sc2.sh = np.random.default_rng(3).normal(0, 2, sc2.sh.shape).astype(np.float32)
    cloud2 = bv.RadFoamGpuCloud(sc2)
    np.testing.assert_array_equal(bv.RadFoamGpuDepthTracer(cloud2).dispatch(cam, (24, 20), start_cell=0).cpu().numpy(),
                                  base)
    with pytest.raises(ValueError):
        bv.RadFoamGpuDepthTracer(cloud).dispatch(cam, (0, 20))
    cloud.deinit()
    cloud2.deinit()


def test_powerfoam_depth_matches_oracle():
    """6,000 PowerFoam sites with Cech adjacency, a 64x48 camera outside the cube."""
    print("case powerfoam depth 64x48", flush=True)
    bv = _bv()
    h = scenes.powerfoam_scene(6000, seed=7, sh_degree=2, density_scale=3.0)
    cloud = bv.RadFoamGpuCloud(h)
    cam = scenes.looking_at((1.2, 0.4, 1.1), (0, 0, 0), np.deg2rad(50), 4 / 3, 100.0)
    tracer = bv.PowerFoamGpuDepthTracer(cloud, bv.RadFoamDepthSettings(), resolution=(64, 48))
    img = tracer.dispatch(cam)
    assert tuple(img.shape) == (48, 64, 4)
    o, d = oracle.camera_rays(cam.position, cam.orientation, cam.tan_half(), cam.principal, 64, 48)
    ref = oracle.trace_splats(oracle.of_host(h), o, d, 0, depth=100.0, eval_mode=bv.EVAL_OPACITY)
    _check(img, ref, "powerfoam 64x48")
    assert ref["rgba"][:, 3].mean() > 0.2
    # the splat map's alpha is the SH splat render's alpha: the same supports and densities
    rgba = bv.PowerFoamGpuSplatTracer(cloud).dispatch(cam, (64, 48))["rgba"].cpu().numpy()
    assert float(np.abs(rgba[:, 3] - img.cpu().numpy().reshape(-1, 4)[:, 1]).max()) <= 1e-6
    cloud.deinit()


def test_powerfoam_depth_rejects_bad_setup(c1_scene):
    bv = _bv()
    walk_cloud = bv.RadFoamGpuCloud(c1_scene)
    with pytest.raises(ValueError):
        bv.PowerFoamGpuDepthTracer(walk_cloud, resolution=(8, 8))
    h = scenes.powerfoam_scene(256, seed=1)
    cloud = bv.RadFoamGpuCloud(h)
    with pytest.raises(ValueError):
        bv.PowerFoamGpuDepthTracer(cloud, resolution=(0, 8))
    walk_cloud.deinit()
    cloud.deinit()
