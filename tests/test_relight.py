"""Point-light shading and the relit walk in the oracle, pinned by closed forms.
CPU only."""
import math

import numpy as np

import oracle


def test_pbr_point_light_closed_form():
    """tests/relight_gpu_vs_cpu.rs:412-470 inputs.  At normal incidence
    (n = v = l) GGX reduces to D = 1/(pi a^2), G = 1, F = f0, so the response is
    radiance * (albedo + f0 / (4 a^2)) with a = roughness^2."""
    light = dict(position=[0, 0, -4], direction=[0, 0, 1], intensity=[16, 8, 4], exponent=2.0)
    got = oracle.shade_point_light(light, [0, 0, 0], [0, 0, -1], [0, 0, -1], [0.5, 0.25, 0.75], 0.5)
    radiance = np.array([16, 8, 4]) / 16.0
    a2 = (0.5 ** 2) ** 2
    np.testing.assert_allclose(got, radiance * (np.array([0.5, 0.25, 0.75]) + 0.04 / (4 * a2)), rtol=1e-6)


def test_point_light_lambert_falloff_and_rejections():
    albedo = [0.5, 0.5, 0.5]
    # isotropic light, tilted normal: albedo * I / d^2 * cos
    light = dict(position=[0, 0, -2], intensity=[8, 8, 8], exponent=0.0)
    n = [0, math.sin(0.3), -math.cos(0.3)]
    got = oracle.shade_point_light(light, [0, 0, 0], n, [0, 0, -1], albedo, 1.0, f0=(0, 0, 0))
    np.testing.assert_allclose(got, 0.5 * 8 / 4 * math.cos(0.3), rtol=1e-6)
    # angular falloff max(dot(direction, away), 0)^exponent, away = (point - light) / |.|
    spot = dict(position=[0, 0, -2], direction=[0, math.sin(0.4), math.cos(0.4)], intensity=[8, 8, 8], exponent=3.0)
    got = oracle.shade_point_light(spot, [0, 0, 0], [0, 0, -1], [0, 0, -1], albedo, 1.0, f0=(0, 0, 0))
    np.testing.assert_allclose(got, 0.5 * 8 / 4 * math.cos(0.4) ** 3, rtol=1e-5)
    # a light behind the surface or at the point contributes nothing
    behind = dict(position=[0, 0, 2], intensity=[8, 8, 8], exponent=0.0)
    assert np.all(oracle.shade_point_light(behind, [0, 0, 0], [0, 0, -1], [0, 0, -1], albedo, 1.0) == 0)
    assert np.all(oracle.shade_point_light(light, [0, 0, -2], [0, 0, -1], [0, 0, -1], albedo, 1.0) == 0)


def test_relit_single_power_cell_closed_form():
    """One power cell (radius 1, density s) seen from z=-2: the integrated
    interval is [1, 3], its entry point (0,0,-1) has normal -z, so one
    isotropic light at (0,0,-5) gives alpha * albedo * I / 16; with
    visibility the shadow walk from the depth point (the origin) leaves the
    support sphere after length 1, scaling by exp(-s)."""
    s = 1.5
    sc = oracle.Scene(np.zeros((1, 3), np.float32), np.array([s], np.float32), np.zeros((1, 3), np.float32), 0,
                      np.array([0, 0], np.uint32), np.zeros(0, np.uint32), np.array([1.0], np.float32))
    mat = np.array([[0.2, 0.4, 0.8, 1.0]], np.float32)
    light = dict(position=[0, 0, -5], intensity=[16, 16, 16], exponent=0.0)
    alpha = 1 - math.exp(-2 * s)
    for vis in (False, True):
        rgba = oracle.trace_relight(sc, mat, [[0, 0, -2]], [[0, 0, 1]], 0, [light], visibility=vis,
                                    specular_f0=(0, 0, 0), weight_threshold=0.0, depth=10.0)
        scale = math.exp(-s) if vis else 1.0
        np.testing.assert_allclose(rgba[0, :3], scale * alpha * mat[0, :3] * 16 / 16, rtol=1e-5)
        assert abs(rgba[0, 3] - alpha) < 1e-6


def test_relit_lights_add_linearly():
    """The relit colour is a sum over lights: two lights together equal the sum
    of each alone (visibility on, so the shadow walks are per light too)."""
    from bv import scenes

    hs = scenes.power_scene(3000, seed=9)
    sc = oracle.of_host(hs)
    rng = np.random.default_rng(4)
    L = scenes.random_lights(2, rng)
    o = rng.uniform(-0.5, 0.5, size=(64, 3)).astype(np.float32)
    d = rng.normal(size=(64, 3)).astype(np.float32)
    st = np.array([scenes.containing_cell(hs, x) for x in o], np.uint32)
    kw = dict(visibility=True, depth=5.0)
    both = oracle.trace_relight(sc, hs.material, o, d, st, L, **kw)
    a = oracle.trace_relight(sc, hs.material, o, d, st, L[:1], **kw)
    b = oracle.trace_relight(sc, hs.material, o, d, st, L[1:], **kw)
    np.testing.assert_allclose(both[:, :3], a[:, :3] + b[:, :3], rtol=1e-5, atol=1e-6)
    assert np.all(both[:, 3] == a[:, 3])


def test_per_light_visibility_splits_by_light():
    """A per-light visibility flag (relight.wgsl's per-light shadow toggle)
    with the lighting-wide flag off: each light contributes what it does alone,
    with a shadow walk when its own flag is set and without one otherwise; a
    flagged light alone equals the lighting-wide flag on that light, and the
    shadowed part differs from the unshadowed one."""
    from bv import scenes

    hs = scenes.power_scene(3000, seed=9)
    sc = oracle.of_host(hs)
    rng = np.random.default_rng(5)
    L = scenes.random_lights(3, rng)
    o = rng.uniform(-0.5, 0.5, size=(64, 3)).astype(np.float32)
    d = rng.normal(size=(64, 3)).astype(np.float32)
    st = np.array([scenes.containing_cell(hs, x) for x in o], np.uint32)
    flags = [True, False, True]
    mixed = [dict(light, visibility=f) for light, f in zip(L, flags)]
    got = oracle.trace_relight(sc, hs.material, o, d, st, mixed, visibility=False, depth=5.0)
    parts = [oracle.trace_relight(sc, hs.material, o, d, st, [light], visibility=f, depth=5.0)
             for light, f in zip(L, flags)]
    np.testing.assert_allclose(got[:, :3], sum(p[:, :3] for p in parts), rtol=1e-5, atol=1e-6)
    alone = oracle.trace_relight(sc, hs.material, o, d, st, mixed[:1], visibility=False, depth=5.0)
    np.testing.assert_array_equal(alone, parts[0])
    free = oracle.trace_relight(sc, hs.material, o, d, st, L[:1], visibility=False, depth=5.0)
    assert np.any(np.abs(free[:, :3] - parts[0][:, :3]) > 1e-4), "the shadow walk changed nothing"
    # the lighting-wide flag overrides: every light casts a visibility ray
    every = oracle.trace_relight(sc, hs.material, o, d, st, mixed, visibility=True, depth=5.0)
    np.testing.assert_array_equal(every, oracle.trace_relight(sc, hs.material, o, d, st, L, visibility=True, depth=5.0))
