"""Oracle of the multi-object scene trace (trace_scene,
blade-volume/shaders/scene_trace_core.wgsl:406-465), pinned on the CPU.

The reference has no CPU scene tracer and its own scene tests
(blade-volume/tests/scene_api.rs) cover bookkeeping only, so the pins are
closed forms on the reference's tiny fixture (tests/data/radfoam_tiny_ascii.ply:
a ray along +z finds no exit face and cell 0, density 0.1, absorbs until the
walk ends) plus invariants:
* one object: alpha = 1 - exp(-0.1 (t_far - t_near)) of its bounding sphere
  (integration window and far end);
* two objects: front-to-back compositing, (t_near, index) order with ties to
  the lower index, the early-out between objects;
* a scaled object integrates over the world ray's t (the unnormalised
  object-space direction);
* rotating and translating a scene and its rays together changes nothing
  (a transposed or misapplied matrix fails it);
* an object whose sphere holds the whole walk equals the plain walk
  (trace_one_ray, pinned by the oracle's own tests).
"""
import math
import os

import numpy as np

import oracle
from bv import io, scenes

from conftest import GOLDEN

IDENT = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)


def _tiny(dc=None):
    m = io.load_radfoam_ply(os.path.join(GOLDEN, "radfoam_tiny_ascii.ply"))
    s = oracle.of_host(m)
    if dc is not None:
        sh = np.zeros_like(s.sh)
        sh[:, :3] = (np.asarray(dc, np.float32) - np.float32(0.5)) / np.float32(scenes.SH_C0)
        s = oracle.Scene(s.xyz, s.density, sh, s.sh_degree, s.offsets, s.neighbors, s.radius)
    return s


def _translate(dz):
    m = IDENT.copy()
    m[11] = -dz  # world -> object
    return m


def _ray():
    return np.array([[0.5, 0.5, -1.0]], np.float32), np.array([[0.0, 0.0, 1.0]], np.float32)


def test_one_object_integrates_its_sphere_interval():
    o, d = _ray()
    # sphere centre (0.5, 0.5, 0), radius 0.5: t_near = 0.5, t_far = 1.5
    r = oracle.trace_objects([(_tiny(), IDENT, (0.5, 0.5, 0.0, 0.5), 0)], o, d, eval_mode=oracle.EVAL_CONSTANT,
                             weight_threshold=1e-4)
    a = 1 - math.exp(-0.1 * 1.0)
    np.testing.assert_allclose(r[0], [a, a, a, a], atol=1e-6)
    # origin inside the sphere: t_near clamps to 0 (radius 2: roots -1 and 3)
    r = oracle.trace_objects([(_tiny(), IDENT, (0.5, 0.5, 0.0, 2.0), 0)], o, d, eval_mode=oracle.EVAL_CONSTANT,
                             weight_threshold=1e-4)
    a = 1 - math.exp(-0.1 * 3.0)
    np.testing.assert_allclose(r[0], [a, a, a, a], atol=1e-6)
    # a missed sphere contributes nothing
    r = oracle.trace_objects([(_tiny(), IDENT, (5.0, 5.0, 0.0, 0.5), 0)], o, d)
    assert np.all(r == 0)


def test_two_objects_composite_front_to_back_in_entry_order():
    o, d = _ray()
    a = 1 - math.exp(-0.1)
    red, blue = _tiny((1.0, 0.0, 0.0)), _tiny((0.0, 0.0, 1.0))
    near = (red, IDENT, (0.5, 0.5, 0.0, 0.5), 0)
    far = (blue, _translate(10.0), (0.5, 0.5, 10.0, 0.5), 0)
    want = [a, 0.0, (1 - a) * a, 1 - (1 - a) ** 2]
    for objs in ([near, far], [far, near]):
        r = oracle.trace_objects(objs, o, d, weight_threshold=1e-4)
        np.testing.assert_allclose(r[0], want, atol=1e-6)
    # equal entry depth: the lower index is visited first (object_hit_less)
    same_r = (red, IDENT, (0.5, 0.5, 0.0, 0.5), 0)
    same_b = (blue, IDENT, (0.5, 0.5, 0.0, 0.5), 0)
    r = oracle.trace_objects([same_b, same_r], o, d, weight_threshold=1e-4)
    np.testing.assert_allclose(r[0], [(1 - a) * a, 0.0, a, 1 - (1 - a) ** 2], atol=1e-6)
    # transmittance below the threshold after the first object ends the ray
    r = oracle.trace_objects([near, far], o, d, weight_threshold=0.95)
    np.testing.assert_allclose(r[0], [a, 0.0, 0.0, a], atol=1e-6)


def test_scaled_object_integrates_over_world_t():
    # object scaled by 2: world sphere centre (1, 1, 0), radius 1; the ray from
    # (1, 1, -2) along +z is inside it for t in [1, 3], so alpha = 1 - exp(-0.2)
    m = np.array([0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0], np.float32)
    o = np.array([[1.0, 1.0, -2.0]], np.float32)
    d = np.array([[0.0, 0.0, 1.0]], np.float32)
    r = oracle.trace_objects([(_tiny(), m, (1.0, 1.0, 0.0, 1.0), 0)], o, d, eval_mode=oracle.EVAL_CONSTANT,
                             weight_threshold=1e-4)
    a = 1 - math.exp(-0.2)
    np.testing.assert_allclose(r[0], [a, a, a, a], atol=1e-6)


def _rot(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * k + (1 - math.cos(ang)) * (k @ k)


def test_rigid_motion_of_scene_and_rays_changes_nothing(small_scene):
    sc = oracle.of_host(small_scene)
    rng = np.random.default_rng(4)
    n = 64
    cam = np.array([0.05, -0.02, 0.03])
    o = np.tile(cam, (n, 1))
    d = rng.normal(size=(n, 3))
    start = oracle.containing_cell(sc, cam.astype(np.float32))
    base = oracle.trace_objects([(sc, IDENT, (0.0, 0.0, 0.0, 2.0), start)], o, d, dtype=np.float64)
    R = _rot([0.3, -0.5, 0.8], 0.9)
    t = np.array([3.0, -1.0, 2.0])
    w2o = np.concatenate([R.T, (-R.T @ t)[:, None]], axis=1)  # object = R^T (world - t)
    moved = oracle.trace_objects([(sc, w2o.reshape(12), (*t, 2.0), start)], (o @ R.T + t), (d @ R.T),
                                 dtype=np.float64)
    np.testing.assert_allclose(moved, base, atol=1e-9)
    # the transposed matrix is a different scene
    wrong = np.concatenate([R, (-R @ t)[:, None]], axis=1)
    bad = oracle.trace_objects([(sc, wrong.reshape(12), (*t, 2.0), start)], (o @ R.T + t), (d @ R.T),
                               dtype=np.float64)
    assert np.abs(bad - base).max() > 1e-3


def test_enclosing_sphere_equals_the_plain_walk(small_scene):
    sc = oracle.of_host(small_scene)
    rng = np.random.default_rng(8)
    n = 128
    cam = np.array([0.1, 0.2, -0.1], np.float32)
    o = np.tile(cam, (n, 1)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    start = oracle.containing_cell(sc, cam)
    # centre at the camera, radius 50: t_near = 0, t_far = 50 for every ray
    r = oracle.trace_objects([(sc, IDENT, (*cam, 50.0), start)], o, d)
    ref = oracle.trace_rays(sc, o, d, start, depth=50.0)["rgba"]
    np.testing.assert_allclose(r, ref, atol=2e-6)
    # axis-aligned rays normalise exactly: bit-identical
    ax = np.array([[1, 0, 0], [0, -1, 0], [0, 0, 1]], np.float32)
    oa = np.tile(cam, (3, 1))
    r = oracle.trace_objects([(sc, IDENT, (*cam, 50.0), start)], oa, ax)
    assert r.tobytes() == oracle.trace_rays(sc, oa, ax, start, depth=50.0)["rgba"].tobytes()
