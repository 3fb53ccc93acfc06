"""Device Cech adjacency builder (bv_compute_cech) against adjacency::compute_cech
(adjacency.rs:406-540, rows sorted by build_symmetric_csr :55-85).

Bar: the CSR is identical (offsets and every neighbour) to an O(N^2) FP32
brute force of the reference predicate |p_i - p_j|^2 <= (r_i + r_j)^2 at small
sizes, and to the host k-d tree builder (scenes.cech_csr, itself pinned to the
brute force in test_splat_oracle.py) at the serving scene's smallest size.  At
2M sites, where neither finishes in seconds, the rows are checked for the
properties the reference's graph has by construction (symmetric, ascending,
no self edge) and a sample of rows against the brute force over all sites.
"""
import numpy as np
import pytest

from bv import scenes

pytestmark = pytest.mark.gpu


def _bv():
    import bv

    return bv


def brute_rows(xyz, r, rows=None):
    """Rows of the reference predicate in FP32: ((dx*dx + dy*dy) + dz*dz) <=
    (r_i + r_j)^2, radii clamped below at 0 (NaN -> 0 as Rust's f32::max),
    i != j."""
    xyz = np.asarray(xyz, np.float32)
    r = np.fmax(np.asarray(r, np.float32), np.float32(0))
    out = []
    for i in (range(len(r)) if rows is None else rows):
        d = xyz[i][None, :] - xyz
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        s = r[i] + r
        ok = d2 <= s * s
        ok[i] = False
        out.append(np.nonzero(ok)[0].astype(np.uint32))
    return out


def assert_csr_equals_rows(off, nb, rows):
    assert len(off) == len(rows) + 1 and off[0] == 0
    lens = np.array([len(x) for x in rows], np.int64)
    np.testing.assert_array_equal(off.astype(np.int64), np.concatenate([[0], np.cumsum(lens)]))
    np.testing.assert_array_equal(nb, np.concatenate(rows) if rows else np.zeros(0, np.uint32))


@pytest.mark.parametrize("seed,n,rmax", [(0, 2000, 0.3), (1, 3000, 0.05), (2, 257, 2.0)])
def test_cech_matches_brute_force(seed, n, rmax):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(1e-3), np.log(rmax), n)).astype(np.float32)
    off, nb = _bv().compute_cech(xyz, r)
    assert_csr_equals_rows(off, nb, brute_rows(xyz, r))


def test_cech_edge_cases():
    bv = _bv()
    # no sites and one site: every row empty
    off, nb = bv.compute_cech(np.zeros((0, 3), np.float32), np.zeros(0, np.float32))
    assert list(off) == [0] and len(nb) == 0
    off, nb = bv.compute_cech(np.zeros((1, 3), np.float32), np.ones(1, np.float32))
    assert list(off) == [0, 0] and len(nb) == 0
    # spheres that exactly touch (|p0 - p1| = 0.5 = r0 + r1 in FP32) are adjacent
    off, nb = bv.compute_cech(np.array([[0, 0, 0], [0.5, 0, 0], [2, 0, 0]], np.float32),
                              np.array([0.25, 0.25, 0.25], np.float32))
    assert list(off) == [0, 1, 2, 2] and list(nb) == [1, 0]
    # coincident sites with zero radius (0 <= 0), a negative and a NaN radius
    # (taken as 0), and a radius that reaches everything
    xyz = np.array([[0, 0, 0], [0, 0, 0], [0.1, 0, 0], [0.2, 0, 0], [5, 5, 5], [-3, 1, 0]], np.float32)
    r = np.array([0.0, 0.0, -1.0, np.nan, 100.0, 0.05], np.float32)
    off, nb = bv.compute_cech(xyz, r)
    assert_csr_equals_rows(off, nb, brute_rows(xyz, r))
    assert list(nb[off[0]:off[1]]) == [1, 4]
    # a site at NaN coordinates touches nothing
    xyz = np.array([[0, 0, 0], [np.nan, 0, 0], [0.1, 0, 0]], np.float32)
    off, nb = bv.compute_cech(xyz, np.full(3, 1.0, np.float32))
    assert list(off) == [0, 1, 1, 2] and list(nb) == [2, 0]
    # duplicates of one point in a grid: many identical keys per leaf range
    g = np.stack(np.meshgrid(*[np.arange(6, dtype=np.float32) * 0.1] * 3, indexing="ij"), -1).reshape(-1, 3)
    xyz = np.concatenate([g, g, g])
    r = np.full(len(xyz), 0.05, np.float32)
    off, nb = bv.compute_cech(xyz, r)
    assert_csr_equals_rows(off, nb, brute_rows(xyz, r))
    with pytest.raises(ValueError):
        bv.compute_cech(xyz, r[:-1])


def test_cech_matches_host_builder_on_serving_scene():
    h = scenes.powerfoam_scene(65536, seed=4, adjacency=False)
    off_h, nb_h = scenes.cech_csr(h.xyz, h.radius)
    off, nb = _bv().compute_cech(h.xyz, h.radius)
    np.testing.assert_array_equal(off, off_h)
    np.testing.assert_array_equal(nb, nb_h)


def test_cech_properties_at_2m_sites():
    n = 2_097_152
    h = scenes.powerfoam_scene(n, seed=7, adjacency=False)
    off, nb = _bv().compute_cech(h.xyz, h.radius)
    assert len(off) == n + 1 and off[0] == 0 and off[-1] == len(nb)
    lens = np.diff(off.astype(np.int64))
    assert (lens >= 0).all()
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    cols = nb.astype(np.int64)
    assert (rows != cols).all()
    # ascending within each row (a row boundary may step down)
    step = np.diff(cols)
    same_row = rows[1:] == rows[:-1]
    assert (step[same_row] > 0).all()
    # symmetric: the transposed edge list, sorted, is the edge list itself
    key = rows * n + cols
    tkey = np.sort(cols * n + rows)
    np.testing.assert_array_equal(key, tkey)
    # a sample of rows against the brute force over every site
    rng = np.random.default_rng(0)
    sample = rng.choice(n, 24, replace=False)
    sample = np.concatenate([sample, [int(np.argmax(lens)), int(np.argmax(h.radius))]])
    for i, row in zip(sample, brute_rows(h.xyz, h.radius, sample)):
        np.testing.assert_array_equal(nb[off[i]:off[i + 1]], row)
