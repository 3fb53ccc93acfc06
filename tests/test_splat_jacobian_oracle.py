"""The oracle of the PowerFoam recorder's Jacobian streams (bv_oracle.cpp,
record_splat_jac / splat_interval_jac).

Pins:
  * the reference's two finite-difference fixtures (trace.rs:2647-2771),
    restated for splat semantics with the camera moved back out of the 4 r
    exclusion along the same ray: three collinear sites of radius 2 (cell 1
    enters through the face with 0 and leaves through the face with 2) and a
    single sphere, whose chord 2 sqrt(r^2 - |perp|^2) has the closed-form
    derivative (2 perp / root, 2 r / root);
  * FP64 central differences (Richardson, relative 1e-6) of the clipped dt
    with respect to the (xyz, radius) of the previous, current and next sites
    on a random Cech scene, for every recorded interval whose face choice a
    small step keeps;
  * the recorded cells and dt equal the interval recorder without Jacobians
    (pinned separately); the tangent equals sum_role J . (p_role - o, r_role);
    past max_path_dt the dt clamps and the derivatives vanish; truncation sets
    bit 31 exactly when an interval beyond L exists.
"""
import math

import numpy as np
import pytest

import oracle
from bv import scenes


def _scene(points, radii, rows):
    off, nb = scenes.csr_from_rows(rows)
    n = len(points)
    return oracle.Scene(np.asarray(points, np.float32), np.ones(n, np.float32), np.zeros((n, 3), np.float32), 0,
                        off, nb, np.asarray(radii, np.float32))


def _fd(sc64, o, d, cell, param_cell, comp, depth, h):
    def dt_at(step):
        xyz = sc64.xyz.copy()
        r = sc64.radius.copy()
        if comp < 3:
            xyz[param_cell, comp] += step
        else:
            r[param_cell] += step
        s2 = oracle.Scene(xyz, sc64.density, sc64.sh, sc64.sh_degree, sc64.offsets, sc64.neighbors, r)
        iv = oracle.splat_interval_jacobian64(s2, o, d, cell, depth=depth)
        return None if iv is None else (iv["dt"], iv["previous"], iv["next"])

    vals = [dt_at(h), dt_at(-h), dt_at(h / 2), dt_at(-h / 2)]
    if any(v is None for v in vals):
        return None
    if len({(v[1], v[2]) for v in vals}) != 1:
        return None  # the step changes which face bounds the interval
    f1 = (vals[0][0] - vals[1][0]) / (2 * h)
    f2 = (vals[2][0] - vals[3][0]) / h
    return (4 * f2 - f1) / 3, f1, f2


def _check_interval(sc, o, d, cell, depth=10.0, h=1e-5):
    sc64 = sc.as_dtype(np.float64)
    iv = oracle.splat_interval_jacobian64(sc64, o, d, cell, depth=depth)
    assert iv is not None
    checked = 0
    for role, pc in (("dt_grad_previous", iv["previous"]), ("dt_grad_current", cell), ("dt_grad_next", iv["next"])):
        if role != "dt_grad_current" and pc == cell:
            assert not np.any(iv[role])
            continue
        for comp in range(4):
            r = _fd(sc64, o, d, cell, pc, comp, depth, h)
            if r is None:
                continue
            est, f1, f2 = r
            want = iv[role][comp] + (iv["dt_grad_current"][comp] if role != "dt_grad_current" and pc == cell else 0)
            assert abs(want - est) <= 1e-6 * max(abs(est), abs(want)) + 1e-9, (role, comp, want, est)
            checked += 1
    return iv, checked


def test_reference_three_site_face_fixture():
    # trace.rs:2647-2722: sites (0,0,0), (1,0,0), (2,0,0), radius 2, chain adjacency
    sc = _scene([[0, 0, 0], [1, 0, 0], [2, 0, 0]], [2, 2, 2], [[1], [0, 2], [1]])
    o = np.array([-9.0, 0.2, -0.1])
    d = np.array([1.0, 0.05, 0.02])
    d = d / np.linalg.norm(d)
    iv, checked = _check_interval(sc, o, d, 1, depth=20.0)  # the depth moves back with the camera
    assert iv["previous"] == 0 and iv["next"] == 2
    assert checked == 12


def test_reference_single_sphere_fixture_closed_form():
    # trace.rs:2724-2771: one site of radius 1, the ray along +z
    sc = _scene([[0, 0, 0]], [1], [[]])
    o = np.array([0.2, 0.1, -5.0])
    d = np.array([0.0, 0.0, 1.0])
    iv, checked = _check_interval(sc, o, d, 0)
    assert iv["previous"] == 0 and iv["next"] == 0
    root = math.sqrt(1.0 - 0.05)
    np.testing.assert_allclose(iv["dt"], 2 * root, rtol=1e-12)
    np.testing.assert_allclose(iv["dt_grad_current"], [0.4 / root, 0.2 / root, 0.0, 2.0 / root], rtol=1e-12,
                               atol=1e-15)
    assert checked == 4


def _cech_scene(n=600, seed=3):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(0.05), np.log(0.25), n)).astype(np.float32)
    off, nb = scenes.cech_csr(xyz, r)
    return oracle.Scene(xyz, np.exp(rng.normal(size=n)).astype(np.float32), np.zeros((n, 3), np.float32), 0, off,
                        nb, r)


def test_recorded_jacobians_match_fp64_differences():
    sc = _cech_scene()
    rng = np.random.default_rng(4)
    o = np.tile(np.array([[0.0, 0.0, -3.0]], np.float32), (12, 1))
    d = np.column_stack([rng.uniform(-0.3, 0.3, 12), rng.uniform(-0.3, 0.3, 12), np.ones(12)]).astype(np.float32)
    rec = oracle.record_splat_jacobians(sc, o, d, max_steps=64, depth=10.0)
    total = 0
    for i in range(12):
        cnt = int(rec["status"][i] & 0x7FFFFFFF)
        dd = d[i].astype(np.float64) / np.linalg.norm(d[i].astype(np.float64))
        for k in range(0, cnt, 3):
            _, checked = _check_interval(sc, o[i].astype(np.float64), dd, int(rec["cells"][i, k]))
            total += checked
    assert total >= 200


def test_rows_tangent_clamp_and_truncation():
    sc = _cech_scene(seed=5)
    rng = np.random.default_rng(6)
    o = rng.uniform(-0.2, 0.2, size=(16, 3)).astype(np.float32) + np.array([0, 0, -3], np.float32)
    d = np.column_stack([rng.uniform(-0.3, 0.3, 16), rng.uniform(-0.3, 0.3, 16), np.ones(16)]).astype(np.float32)
    L = 48
    rec = oracle.record_splat_jacobians(sc, o, d, max_steps=L, depth=10.0)
    plain = oracle.record_splats(sc, o, d, max_steps=L, depth=10.0)
    cnt = rec["status"] & 0x7FFFFFFF
    assert np.array_equal(cnt, plain["count"])
    live = np.arange(L)[None, :] < cnt[:, None]
    assert np.array_equal(rec["cells"][live], plain["cells"][live])
    assert np.array_equal(rec["dts"][live], plain["dts"][live])
    # tangent = sum over roles of J . (site - origin, radius), in FP32
    xyz, r = sc.xyz, sc.radius
    for i, k in zip(*np.nonzero(live)):
        t = 0.0
        for role, ids in (("dt_grad_previous", "previous_cells"), ("dt_grad_current", "cells"),
                          ("dt_grad_next", "next_cells")):
            c = int(rec[ids][i, k])
            g = np.append(xyz[c] - o[i], r[c]).astype(np.float64)
            t += float(np.dot(rec[role][i, k].astype(np.float64), g))
        assert abs(t - rec["dt_reference_tangent"][i, k]) <= 1e-5 * max(1.0, abs(t))
    # truncation: bit 31 exactly when the full list is longer than L
    full = oracle.record_splats(sc, o, d, max_steps=4096, depth=10.0)
    short = oracle.record_splat_jacobians(sc, o, d, max_steps=5, depth=10.0)
    assert np.array_equal((short["status"] >> 31) == 1, full["count"] > 5)
    # past max_path_dt the dt clamps and its derivatives vanish
    cl = oracle.record_splat_jacobians(sc, o, d, max_steps=L, depth=10.0, max_path_dt=0.02)
    big = live & (rec["dts"] > 0.02)
    assert big.any()
    assert np.all(cl["dts"][big] == np.float32(0.02))
    for role in ("dt_grad_previous", "dt_grad_current", "dt_grad_next"):
        assert np.all(cl[role][big] == 0)
    assert np.all(cl["dt_reference_tangent"][big] == 0)
