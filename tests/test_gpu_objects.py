"""Multi-object scene trace on the device (bv_trace_objects, bv.Scene) against
the oracle (trace_scene, scene_trace_core.wgsl:406-465; Scene,
blade-volume/src/scene.rs).

Bar: within 1e-5 absolute in FP32 of the oracle's composition on the same
world rays, objects and start cells (the same walk as the single-scene trace,
whose colour bar is 1e-5: libm expf against the device exp), for
* one object with the identity transform (also against the plain trace);
* three objects, translated, rotated and uniformly scaled, with overlapping
  bounding spheres, through the camera entry (Scene.dispatch);
* PowerFoam objects (the sphere-clipped walk with a non-unit direction);
* the closed-form two-object composite on the reference's tiny fixture.
Argument errors (too many objects, mixed SH degrees, bad start cells) raise.
"""
import math
import os

import numpy as np
import pytest

import oracle
from bv import io, scenes
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _bv():
    import bv

    return bv


def _objects_for_oracle(scene):
    bv = _bv()
    out = []
    for k, ob in enumerate(scene._objects):
        c, r = scene.bounds(k)
        out.append((oracle.of_host(ob["cloud"].host), ob["transform"].world_to_object().reshape(12),
                    (float(c[0]), float(c[1]), float(c[2]), float(r)), ob["start"]))
    assert len(out) == scene.object_count() and bv is not None
    return out


def _rays(n, cam, seed):
    rng = np.random.default_rng(seed)
    o = np.tile(np.asarray(cam, np.float32), (n, 1))
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return o, d


def test_single_identity_object_matches_oracle_and_plain_trace(small_scene):
    import torch

    bv = _bv()
    cloud = bv.RadFoamGpuCloud(small_scene)
    sc = bv.Scene()
    sc.add_radfoam(cloud)
    cam = (0.1, -0.2, 0.05)
    sc.update_camera(cam)
    o, d = _rays(4096, cam, 1)
    got = sc.trace_rays(torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()).cpu().numpy()
    want = oracle.trace_objects(_objects_for_oracle(sc), o, d)
    err = float(np.abs(got - want).max())
    print(f"single object: max |rgba - oracle| = {err:.3g}", flush=True)
    assert err <= TOL
    # the bounding sphere holds the whole scene and the camera: the plain walk to t_far
    c, r = sc.bounds(0)
    start = sc._objects[0]["start"]
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    oc = o - c
    b = (oc * dn).sum(1)
    t_far = -b + np.sqrt(b * b - ((oc * oc).sum(1) - r * r))
    ref = np.stack([oracle.trace_rays(oracle.of_host(small_scene), o[i:i + 1], d[i:i + 1], start,
                                      depth=float(t_far[i]))["rgba"][0] for i in range(64)])
    assert float(np.abs(got[:64] - ref).max()) <= TOL


def test_three_transformed_objects_camera_frame(small_scene):
    bv = _bv()
    rng = np.random.default_rng(21)
    clouds = []
    for seed in (11, 12, 13):
        r2 = np.random.default_rng(seed)
        n = 384
        xyz = r2.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        off, nb = scenes.delaunay_csr(xyz)
        dens = np.exp(r2.normal(size=n)).astype(np.float32)
        sh = r2.normal(scale=0.3, size=(n, 48)).astype(np.float32)
        clouds.append(bv.RadFoamGpuCloud(scenes.HostScene(xyz, dens, sh, 3, off, nb)))
    sc = bv.Scene()
    sc.add_radfoam(clouds[0], bv.Transform(position=(0.0, 0.0, 0.0)))
    q = rng.normal(size=4)
    sc.add_radfoam(clouds[1], bv.Transform(position=(1.2, 0.3, -0.4), rotation=tuple(q / np.linalg.norm(q))))
    sc.add_radfoam(clouds[2], bv.Transform(position=(-1.0, -0.5, 0.8), scale=(0.6, 0.6, 0.6)))
    cam = scenes.looking_at(np.array([0.0, 0.5, 4.0], np.float32), np.zeros(3, np.float32), math.radians(60), 1.0,
                            100.0)
    w = h = 64
    img = sc.dispatch(cam, (w, h)).cpu().numpy().reshape(h * w, 4)
    o, d = bv.camera_rays(cam, w, h)
    want = oracle.trace_objects(_objects_for_oracle(sc), o.cpu().numpy(), d.cpu().numpy())
    err = float(np.abs(img - want).max())
    print(f"three objects: max |rgba - oracle| = {err:.3g}, covered {float((want[:, 3] > 0).mean()):.2f}", flush=True)
    assert (want[:, 3] > 0).mean() > 0.2
    assert err <= TOL


def test_powerfoam_objects():
    import torch

    bv = _bv()
    pf = scenes.powerfoam_scene(2048, seed=5)
    cloud = bv.RadFoamGpuCloud(pf)
    sc = bv.Scene()
    sc.add_radfoam(cloud, bv.Transform(position=(0.2, 0.0, 0.0), scale=(1.5, 1.5, 1.5)))
    sc.add_radfoam(cloud, bv.Transform(position=(-2.5, 0.1, -0.3)))
    cam = (0.0, 0.1, 3.0)
    sc.update_camera(cam)
    rng = np.random.default_rng(3)
    o = np.tile(np.asarray(cam, np.float32), (2048, 1))
    # toward points scattered around the two objects' centres
    centres = np.array([[0.2, 0.0, 0.0], [-2.5, 0.1, -0.3]])[rng.integers(0, 2, 2048)]
    d = (centres + rng.normal(scale=0.3, size=(2048, 3)) - o).astype(np.float32)
    got = sc.trace_rays(torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()).cpu().numpy()
    want = oracle.trace_objects(_objects_for_oracle(sc), o, d)
    err = float(np.abs(got - want).max())
    print(f"powerfoam objects: max |rgba - oracle| = {err:.3g}", flush=True)
    assert (want[:, 3] > 0).mean() > 0.3
    assert err <= TOL


def test_two_object_closed_form_on_the_tiny_fixture():
    import torch

    bv = _bv()
    m = io.load_radfoam_ply(os.path.join(GOLDEN, "radfoam_tiny_ascii.ply"))
    red = scenes.HostScene(m.xyz, m.density, m.sh.copy(), m.sh_degree, m.offsets, m.neighbors)
    blue = scenes.HostScene(m.xyz, m.density, m.sh.copy(), m.sh_degree, m.offsets, m.neighbors)
    red.sh[:] = 0
    blue.sh[:] = 0
    red.sh[:, :3] = (np.array([1.0, 0.0, 0.0], np.float32) - np.float32(0.5)) / np.float32(scenes.SH_C0)
    blue.sh[:, :3] = (np.array([0.0, 0.0, 1.0], np.float32) - np.float32(0.5)) / np.float32(scenes.SH_C0)
    cr, cb = bv.RadFoamGpuCloud(red), bv.RadFoamGpuCloud(blue)
    lib = bv._lib.load()
    objs = (bv._lib.SceneObject * 2)()
    ident = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
    for k, (cl, dz) in enumerate([(cb, 10.0), (cr, 0.0)]):
        objs[k].scene = cl.handle.value
        objs[k].world_to_object[:] = ident[:11] + [-dz]
        objs[k].center[:] = [0.5, 0.5, dz]
        objs[k].radius = 0.5
        objs[k].start_cell = 0
    o = torch.tensor([[0.5, 0.5, -1.0]], device="cuda")
    d = torch.tensor([[0.0, 0.0, 1.0]], device="cuda")
    out = torch.zeros((1, 4), device="cuda")
    rays = bv._rays(o, d, None, 0)
    st = bv.TraceSettings(weight_threshold=1e-4)
    bv.check(lib.bv_trace_objects(objs, 2, bv.C.byref(rays), bv.C.byref(st.c()), out.data_ptr(), None))
    torch.cuda.synchronize()
    a = 1 - math.exp(-0.1)
    np.testing.assert_allclose(out.cpu().numpy()[0], [a, 0.0, (1 - a) * a, 1 - (1 - a) ** 2], atol=1e-6)


def test_argument_errors(small_scene, c1_scene):
    import torch

    bv = _bv()
    a, b = bv.RadFoamGpuCloud(small_scene), bv.RadFoamGpuCloud(c1_scene)
    o = torch.zeros((4, 3), device="cuda")
    d = torch.ones((4, 3), device="cuda")
    sc = bv.Scene()
    sc.add_radfoam(a)
    sc.add_radfoam(b)  # SH degree 3 beside SH degree 0
    with pytest.raises(bv.BvError):
        sc.trace_rays(o, d)
    sc = bv.Scene()
    for _ in range(8):
        sc.add_radfoam(a)
    with pytest.raises(ValueError):
        sc.add_radfoam(a)
    lib = bv._lib.load()
    objs = (bv._lib.SceneObject * 9)()
    for k in range(9):
        objs[k].scene = a.handle.value
        objs[k].world_to_object[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]
        objs[k].radius = 1.0
    rays = bv._rays(o, d, None, 0)
    out = torch.zeros((4, 4), device="cuda")
    st = bv.TraceSettings()
    with pytest.raises(bv.BvError):
        bv.check(lib.bv_trace_objects(objs, 9, bv.C.byref(rays), bv.C.byref(st.c()), out.data_ptr(), None))
    objs[0].start_cell = small_scene.n  # out of range
    with pytest.raises(bv.BvError):
        bv.check(lib.bv_trace_objects(objs, 1, bv.C.byref(rays), bv.C.byref(st.c()), out.data_ptr(), None))
    objs[0].start_cell = 0
    objs[0].world_to_object[:] = [0.0] * 12  # singular
    with pytest.raises(bv.BvError):
        bv.check(lib.bv_trace_objects(objs, 1, bv.C.byref(rays), bv.C.byref(st.c()), out.data_ptr(), None))
