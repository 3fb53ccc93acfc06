"""The batched recorder's ray restatement (oracle.recorder_rays) pinned without
the reference: the recorder's quaternion rotation against scipy's rotation
matrices, against the other rotation formula the reference uses for camera
rays (common.wgsl's qrot, oracle.camera_rays), its pixel-index arithmetic
(row-major layout, image_pixel_offset with u32 wrap-around), and the slicing of
oracle.record_batch over a [capacity, L] buffer.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle
from bv import scenes


def _view(q, W=7, H=5, fov=(1.1, 0.8), principal=(0.0, 0.0), pos=(0.25, -0.5, 1.0), off=0):
    return dict(position=np.asarray(pos, np.float32), orientation=np.asarray(q, np.float32),
                fov=np.asarray(fov, np.float32), principal=np.asarray(principal, np.float32), image_width=W,
                image_height=H, image_pixel_offset=off)


def test_recorder_rotation_matches_rotation_matrix():
    rng = np.random.default_rng(0)
    W, H = 9, 6
    for q in scenes.random_quaternions(16, rng):
        v = _view(q, W, H, principal=(0.1, -0.05))
        o, d = oracle.recorder_rays(v, np.arange(W * H))
        np.testing.assert_array_equal(o, np.broadcast_to(v["position"], o.shape))
        # the unrotated direction: pixel centres in NDC scaled by tan(fov / 2), z = 1
        ix, iy = np.arange(W * H) % W, np.arange(W * H) // W
        th = np.tan(0.5 * v["fov"].astype(np.float64))
        local = np.column_stack([((ix + 0.5) / W * 2 - 1 - 0.1) * th[0], ((iy + 0.5) / H * 2 - 1 + 0.05) * th[1],
                                 np.ones(W * H)])
        want = Rotation.from_quat(q.astype(np.float64)).apply(local)
        want /= np.linalg.norm(want, axis=1, keepdims=True)
        np.testing.assert_allclose(oracle.unit(d), want, atol=2e-6, rtol=0)
        # qrot is quadratic in q: doubling the quaternion scales every term by 4 exactly
        _, d2 = oracle.recorder_rays(_view(2 * q, W, H, principal=(0.1, -0.05)), np.arange(W * H))
        np.testing.assert_array_equal(d2, 4 * d)


def test_recorder_rays_agree_with_camera_rays():
    """Two rotation formulas of the reference (the recorder's and common.wgsl's)
    give the same unit directions to FP32 rounding, pixel for pixel."""
    rng = np.random.default_rng(1)
    W, H = 13, 8
    for q in scenes.random_quaternions(8, rng):
        v = _view(q, W, H, principal=(-0.2, 0.15))
        _, d = oracle.recorder_rays(v, np.arange(W * H))
        _, dc = oracle.camera_rays(v["position"], v["orientation"], oracle.recorder_tan_half(v["fov"]),
                                   v["principal"], W, H)
        np.testing.assert_allclose(oracle.unit(d), dc, atol=1e-6, rtol=0)


def test_identity_orientation_keeps_the_local_vector():
    W, H = 4, 3
    v = _view([0, 0, 0, 1], W, H)
    _, d = oracle.recorder_rays(v, [0, 5, 11])
    th = np.float32(np.tan(np.float64(np.float32(0.5) * v["fov"])))
    assert th[0] == oracle.recorder_tan_half(v["fov"])[0]
    # pixel 5 = (1, 1): x = (1.5 / 4) * 2 - 1 = -0.25, y = (1.5 / 3) * 2 - 1 = 0
    np.testing.assert_array_equal(d[1], np.array([np.float32(-0.25) * th[0], 0.0, 1.0], np.float32))
    # pixels 0 and 11 are the image's opposite corners
    np.testing.assert_array_equal(d[0][:2], -d[2][:2])


def test_pixel_offset_wraps_in_u32():
    v = _view([0.1, 0.2, -0.3, 0.9], 6, 4, off=5)
    idx = np.array([0, 1, 18, 0xFFFFFFFF - 2, 0xFFFFFFFF], np.uint64)
    _, d = oracle.recorder_rays(v, idx.astype(np.uint32))
    v0 = dict(v, image_pixel_offset=0)
    _, want = oracle.recorder_rays(v0, ((idx + 5) % (1 << 32)).astype(np.uint32))
    np.testing.assert_array_equal(d, want)
    # an index past the image keeps the row-major arithmetic (iy >= height), as the shader does
    _, past = oracle.recorder_rays(v0, [6 * 4 + 2])
    _, first = oracle.recorder_rays(v0, [2])
    assert past[0].tolist() != first[0].tolist()


@pytest.fixture(scope="module")
def small_uniform():
    return oracle.of_host(scenes.uniform_scene(2048, seed=5))


def test_record_batch_slices(small_uniform):
    sc = small_uniform
    rng = np.random.default_rng(2)
    cams = [scenes.looking_at(rng.uniform(-0.4, 0.4, 3), rng.uniform(-1, 1, 3), 0.9, 1.25, 100.0) for _ in range(2)]
    views = []
    for c, (off, n, W, H) in zip(cams, ((0, 30, 10, 8), (40, 25, 6, 6))):
        views.append(dict(position=c.position, orientation=c.orientation, fov=c.fov, principal=c.principal,
                          image_width=W, image_height=H, image_pixel_offset=3, pixel_offset=off, num_pixels=n,
                          start_point=oracle.containing_cell(sc, c.position), depth=100.0, max_path_dt=0.05))
    cap, L = 70, 24
    idx = rng.integers(0, 50, cap).astype(np.uint32)
    got = oracle.record_batch(sc, views, idx, capacity=cap, max_steps=L)
    for k, v in enumerate(views):
        a, n = v["pixel_offset"], v["num_pixels"]
        o, d = oracle.recorder_rays(v, idx[a:a + n])
        ref = oracle.record_paths(sc, o, d, v["start_point"], max_steps=L, depth=100.0, max_path_dt=0.05)
        for key in ("cells", "next_cells", "status"):
            np.testing.assert_array_equal(got[key][a:a + n], ref[key])
        np.testing.assert_array_equal(got["dts"][a:a + n], ref["dts"])
        np.testing.assert_array_equal(got["view_index"][a:a + n], k)
        np.testing.assert_array_equal(got["ray_dirs"][a:a + n], oracle.unit(d))
        assert np.all(ref["status"] & 0x7FFFFFFF > 0)
        assert np.all(np.nan_to_num(ref["dts"], nan=0.0) <= np.float32(0.05))
    gap = np.r_[30:40, 65:70]
    assert np.all(got["status"][gap] == 0) and np.all(got["cells"][gap] == oracle.NO_CELL)
    assert np.all(got["view_index"][gap] == oracle.NO_CELL)
