"""Scenes restated from the reference's own tests (test data, not product):

  * ``chain``      tests/radfoam_synth_chain.rs:32-105
  * ``branching``  tests/radfoam_synth_branch.rs:106-277 (defaults :86-99)
  * ``x_grid``     tests/gpu_path_record.rs:45-82
  * ``tetra``      trace.rs:1828-1851 (adjacency from the Delaunay builder)
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from bv.scenes import HostScene, csr_from_rows


def chain(n: int, dz: float, density: float, sh_degree: int, dc=(0.0, 0.0, 0.0)) -> HostScene:
    K = (sh_degree + 1) ** 2
    xyz = np.zeros((n, 3), np.float32)
    xyz[:, 2] = np.arange(n, dtype=np.float32) * np.float32(dz)
    rows = [[j for j in (i - 1, i + 1) if 0 <= j < n] for i in range(n)]
    off, nb = csr_from_rows(rows)
    sh = np.zeros((n, 3 * K), np.float32)
    sh[:, 0:3] = np.asarray(dc, np.float32)
    return HostScene(xyz, np.full(n, density, np.float32), sh, sh_degree, off, nb)


def branching(spine_len=64, dz=0.05, branch_degree=6, branch_radius=0.10, branch_z_offset=0.02, density=0.1,
              sh_degree=3, dc=(0.1, 0.1, 0.1)) -> HostScene:
    n_total = spine_len + spine_len * branch_degree
    K = (sh_degree + 1) ** 2
    pts = np.zeros((n_total, 3), np.float32)
    pts[:spine_len, 2] = np.arange(spine_len, dtype=np.float32) * np.float32(dz)
    angles = [np.float32(j) / np.float32(branch_degree) * np.float32(2 * math.pi) for j in range(branch_degree)]
    for i in range(spine_len):
        for j, a in enumerate(angles):
            k = spine_len + i * branch_degree + j
            pts[k] = (pts[i, 0] + np.float32(branch_radius) * np.cos(a, dtype=np.float32),
                      pts[i, 1] + np.float32(branch_radius) * np.sin(a, dtype=np.float32),
                      pts[i, 2] + np.float32(branch_z_offset))
    rows: List[List[int]] = [[] for _ in range(n_total)]
    for i in range(spine_len):
        if i > 0:
            rows[i].append(i - 1)
        if i + 1 < spine_len:
            rows[i].append(i + 1)
        for j in range(branch_degree):
            b = spine_len + i * branch_degree + j
            rows[i].append(b)
            rows[b].append(i)
    off, nb = csr_from_rows(rows)
    sh = np.zeros((n_total, 3 * K), np.float32)
    sh[:, 0:3] = np.asarray(dc, np.float32)
    if sh_degree >= 1:
        sh[:, 3] = 0.05
    if sh_degree >= 2:
        sh[:, 12] = 0.03
    if sh_degree >= 3:
        sh[:, 27] = 0.02
    return HostScene(pts, np.full(n_total, density, np.float32), sh, sh_degree, off, nb)


def x_grid(n: int) -> HostScene:
    xyz = np.zeros((n, 3), np.float32)
    xyz[:, 0] = np.arange(n, dtype=np.float32)
    rows = [[j for j in (i - 1, i + 1) if 0 <= j < n] for i in range(n)]
    off, nb = csr_from_rows(rows)
    return HostScene(xyz, np.zeros(n, np.float32), np.zeros((n, 3), np.float32), 0, off, nb)


def tetra() -> HostScene:
    xyz = np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0], [0, 0, 0.5]], np.float32)
    rows = [[j for j in range(4) if j != i] for i in range(4)]  # the Delaunay graph of 4 points is complete
    off, nb = csr_from_rows(rows)
    return HostScene(xyz, np.full(4, 5.0, np.float32), np.zeros((4, 3), np.float32), 0, off, nb)
