"""The library's Delaunay builder (bv_compute_delaunay) — CPU only: the builder
runs on the host.

Pins: the edge set equals Qhull's (scipy.spatial.Delaunay, the library behind
the reference's qhull backend, adjacency.rs:564-634) row for row on seeded
general-position sites from 16 to 16,384, uniform and mixture-on-shell; in
general position the Delaunay tetrahedralization is unique, so any exact
builder (the reference's simple_delaunay_lib included) gives these rows.
The CSR satisfies validate_csr_result (adjacency.rs:721-797) at 1,048,576
shell sites.  Degenerate inputs: fewer than 4 sites and coplanar sites are
rejected as AdjacencyError::TooFewPoints / TriangulationFailed; a repeated site
gets an empty row and leaves every other row as without it; the cospherical
corners of a cubic grid get a valid, deterministic triangulation.
"""
import numpy as np
import pytest
from scipy.spatial import Delaunay

import bv
from bv import _lib, scenes


def qhull_csr(xyz):
    return scenes.delaunay_csr(xyz, builder="qhull")


def validate_csr(off, nb, n):
    # adjacency.rs:721-797
    assert off.shape == (n + 1,) and off[0] == 0 and int(off[-1]) == len(nb)
    assert np.all(np.diff(off.astype(np.int64)) >= 0)
    rows = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    assert np.all(nb < n) and not np.any(nb == rows)
    same_row = rows[1:] == rows[:-1]
    assert np.all(nb[1:][same_row] > nb[:-1][same_row]), "rows must be sorted and unique"
    fwd = rows.astype(np.int64) * n + nb
    rev = nb.astype(np.int64) * n + rows
    assert np.array_equal(np.sort(fwd), np.sort(rev)), "every edge needs its reverse"


@pytest.mark.parametrize("n,seed", [(16, 1), (100, 2), (2000, 3), (16384, 0)])
def test_uniform_sites_equal_qhull(n, seed):
    xyz = np.random.default_rng(seed).uniform(-1, 1, size=(n, 3)).astype(np.float32)
    off, nb = bv.compute_delaunay(xyz)
    o2, n2 = qhull_csr(xyz)
    validate_csr(off, nb, n)
    assert np.array_equal(off, o2) and np.array_equal(nb, n2)


def test_shell_sites_equal_qhull():
    xyz = scenes.shell_points(16384, np.random.default_rng(5))
    off, nb = bv.compute_delaunay(xyz)
    o2, n2 = qhull_csr(xyz)
    assert np.array_equal(off, o2) and np.array_equal(nb, n2)


def test_csr_invariants_at_one_million_shell_sites():
    n = 1 << 20
    xyz = scenes.shell_points(n, np.random.default_rng(1))
    off, nb = bv.compute_delaunay(xyz)
    validate_csr(off, nb, n)
    deg = np.diff(off.astype(np.int64))
    assert deg.min() >= 4 and 14.0 < deg.mean() < 17.0  # Poisson-Delaunay mean degree ~15.5


def test_rejects_what_the_reference_rejects():
    with pytest.raises(_lib.BvError, match="at least 4 points"):
        bv.compute_delaunay(np.zeros((3, 3), np.float32))
    flat = np.random.default_rng(0).uniform(-1, 1, size=(50, 3)).astype(np.float32)
    flat[:, 2] = 0.25
    with pytest.raises(_lib.BvError, match="coplanar"):
        bv.compute_delaunay(flat)
    bad = np.zeros((8, 3), np.float32)
    bad[3, 1] = np.nan
    with pytest.raises(_lib.BvError, match="non-finite"):
        bv.compute_delaunay(bad)


def test_repeated_site_gets_an_empty_row():
    xyz = np.random.default_rng(7).uniform(-1, 1, size=(300, 3)).astype(np.float32)
    dup = np.concatenate([xyz, xyz[[17]]])
    off, nb = bv.compute_delaunay(dup)
    validate_csr(off, nb, 301)
    assert off[301] == off[300]  # the repeat's row is empty
    o1, n1 = bv.compute_delaunay(xyz)
    assert np.array_equal(off[:301], o1) and np.array_equal(nb, n1)


def test_cospherical_grid_is_valid_and_deterministic():
    g = np.stack(np.meshgrid(*[np.arange(5, dtype=np.float32)] * 3, indexing="ij"), -1).reshape(-1, 3)
    a = bv.compute_delaunay(g)
    b = bv.compute_delaunay(g)
    validate_csr(*a, len(g))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # every grid edge of the cube lattice is a Delaunay edge in any resolution of the ties
    off, nb = a
    d = np.abs(g[nb] - g[np.repeat(np.arange(len(g)), np.diff(off.astype(np.int64)))]).sum(1)
    lattice = {(i, j) for i in range(len(g)) for j in range(len(g)) if np.abs(g[i] - g[j]).sum() == 1}
    got = set(zip(np.repeat(np.arange(len(g)), np.diff(off.astype(np.int64))).tolist(), nb.tolist()))
    assert lattice <= got
    assert d.max() <= 3.0  # no edge longer than a cube diagonal
