"""Synthetic scenes, cameras and fixtures for the cell walk.

Everything is generated from ``numpy.random.default_rng(seed)`` so every run of
every consumer sees the same bytes.  The logical layout is the reference's
``PointCloudModel`` (blade-volume/src/lib.rs:173-239): positions [N,3] f32,
density [N] f32, SH coefficients [N, 3K] f32 with RGB interleaved per component
(trace.rs:271-272), optional radii [N] f32, CSR adjacency offsets [N+1] u32 /
neighbours [E] u32 with each row sorted ascending and self-free
(adjacency.rs:55-85).

Configuration scenes (sizes and distributions as stated in the workload list):
  * ``uniform_scene``  16,384 sites U[-1,1]^3, SH degree 0
  * ``shell_scene``    1,048,576 sites, 8-component mixture on the unit shell + 10% background, SH degree 3
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

SH_C0 = 0.2820948  # trace.rs:244


@dataclass
class HostScene:
    xyz: np.ndarray                 # [N,3] f32
    density: np.ndarray             # [N] f32
    sh: np.ndarray                  # [N,3K] f32
    sh_degree: int
    offsets: np.ndarray             # [N+1] u32
    neighbors: np.ndarray           # [E] u32
    radius: Optional[np.ndarray] = None
    material: Optional[np.ndarray] = None  # [N,4] albedo rgb + GGX roughness (relighting)

    @property
    def n(self) -> int:
        return int(self.density.shape[0])

    def validate(self) -> None:
        """CSR and layout checks of PointCloudModel::validate (lib.rs:246-483) and
        validate_csr_result (adjacency.rs:721-797)."""
        n = self.n
        K = (min(self.sh_degree, 3) + 1) ** 2
        assert self.xyz.shape == (n, 3) and self.xyz.dtype == np.float32
        assert self.density.dtype == np.float32
        assert self.sh.shape == (n, 3 * K) and self.sh.dtype == np.float32
        assert self.offsets.shape == (n + 1,) and self.offsets.dtype == np.uint32
        assert self.offsets[0] == 0 and int(self.offsets[-1]) == self.neighbors.shape[0]
        assert np.all(np.diff(self.offsets.astype(np.int64)) >= 0)
        assert self.neighbors.dtype == np.uint32
        if self.neighbors.size:
            assert int(self.neighbors.max()) < n
        if self.radius is not None:
            assert self.radius.shape == (n,) and self.radius.dtype == np.float32
        if self.material is not None:
            assert self.material.shape == (n, 4) and self.material.dtype == np.float32


@dataclass
class Camera:
    """CameraParams (camera.rs:9-18): position, depth, orientation (x,y,z,w),
    fov (x, y) in radians and the principal point in NDC."""
    position: np.ndarray
    depth: float
    orientation: np.ndarray
    fov: np.ndarray
    principal: np.ndarray = field(default_factory=lambda: np.zeros(2, np.float32))

    def tan_half(self) -> np.ndarray:
        # tan of the FP32 half angle in double, rounded once to FP32: the same
        # definition the device library uses, so rays agree bit for bit.
        h = np.float32(0.5) * np.asarray(self.fov, np.float32)
        return np.array([math.tan(float(h[0])), math.tan(float(h[1]))], np.float32)


def _quat_from_cols(right, down, forward) -> np.ndarray:
    """Rotation matrix with the given columns -> unit quaternion (x, y, z, w)."""
    m = np.stack([right, down, forward], axis=1).astype(np.float64)
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        w, x, y, z = 0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        w, x, y, z = (m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s
    elif m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        w, x, y, z = (m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s
    else:
        s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        w, x, y, z = (m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    return (q / np.linalg.norm(q)).astype(np.float32)


def orientation_looking(forward, world_up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """camera.rs:27-45: local frame right, down, forward."""
    f = np.asarray(forward, np.float64)
    nf = np.linalg.norm(f)
    if nf == 0:
        return np.array([0, 0, 0, 1], np.float32)
    f = f / nf
    up = np.asarray(world_up, np.float64)
    if abs(f @ up) > 0.999:
        up = np.array([0.0, 0.0, 1.0]) if abs(up[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    down = f * (f @ up) - up
    down /= np.linalg.norm(down)
    right = np.cross(down, f)
    right /= np.linalg.norm(right)
    return _quat_from_cols(right, down, f)


def looking_at(position, target, fov_y: float, aspect: float, depth: float) -> Camera:
    """CameraParams::looking_at, camera.rs:53-68."""
    p = np.asarray(position, np.float32)
    q = orientation_looking(np.asarray(target, np.float64) - p.astype(np.float64))
    fov_x = 2.0 * math.atan(math.tan(0.5 * fov_y) * aspect)
    return Camera(p, float(depth), q, np.array([fov_x, fov_y], np.float32))


def ring_cameras(count: int, radius: float, fov_y: float, aspect: float, depth: float) -> List[Camera]:
    """Cameras on a ring in the y = 0 plane at equal azimuth steps, looking at the origin."""
    cams = []
    for k in range(count):
        a = 2.0 * math.pi * k / count
        cams.append(looking_at((radius * math.cos(a), 0.0, radius * math.sin(a)), (0, 0, 0), fov_y, aspect, depth))
    return cams


# ---------------------------------------------------------------- adjacency

def csr_from_rows(rows: List[List[int]]) -> tuple:
    offsets = np.zeros(len(rows) + 1, np.uint32)
    flat: List[int] = []
    for i, r in enumerate(rows):
        r = sorted(set(int(x) for x in r if int(x) != i))
        flat.extend(r)
        offsets[i + 1] = len(flat)
    return offsets, np.asarray(flat, np.uint32)


def delaunay_csr(xyz: np.ndarray, builder: str = "native") -> tuple:
    """Delaunay edge graph as CSR with sorted, self-free rows (adjacency.rs:55-85,
    142-197).  builder "native": the library's exact incremental builder
    (bv_compute_delaunay); "qhull": scipy's Qhull, the library behind the
    reference's optional qhull backend (adjacency.rs:564-634)."""
    if builder == "native":
        from . import compute_delaunay

        return compute_delaunay(np.asarray(xyz, np.float32))
    from scipy.spatial import Delaunay

    tri = Delaunay(np.asarray(xyz, np.float64))
    indptr, ind = tri.vertex_neighbor_vertices
    n = xyz.shape[0]
    indptr = indptr.astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr))
    order = np.lexsort((ind.astype(np.int64), rows))
    neighbors = ind[order].astype(np.uint32)
    offsets = indptr.astype(np.uint32)
    if np.any(np.diff(indptr) == 0):
        raise ValueError("Delaunay left a site without neighbours (coplanar/duplicate input)")
    return offsets, neighbors


# ---------------------------------------------------------------- configuration scenes

def uniform_scene(n: int = 16_384, seed: int = 0) -> HostScene:
    """n sites U[-1,1]^3; density exp(N(0,1)); SH degree 0 with RGB U[0,1]
    stored as DC = (rgb - 0.5) / C0 (trace.rs:779-782)."""
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    density = np.exp(rng.normal(size=n)).astype(np.float32)
    rgb = rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32)
    sh = ((rgb - np.float32(0.5)) / np.float32(SH_C0)).astype(np.float32)
    off, nb = delaunay_csr(xyz)
    return HostScene(xyz, density, sh, 0, off, nb)


def shell_points(n: int, rng, angular_sigma: float = 0.5) -> np.ndarray:
    """90% from an 8-component mixture on the unit sphere (component centres
    uniform on the sphere, each sample centre + N(0, angular_sigma^2 I)
    re-projected to the sphere, then radius N(1, 0.05^2)), 10% uniform in
    [-1.5, 1.5]^3."""
    nb = n // 10
    nm = n - nb
    c = rng.normal(size=(8, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    k = rng.integers(0, 8, nm)
    u = c[k] + rng.normal(scale=angular_sigma, size=(nm, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = 1.0 + rng.normal(scale=0.05, size=(nm, 1))
    return np.concatenate([u * r, rng.uniform(-1.5, 1.5, size=(nb, 3))]).astype(np.float32)


def shell_scene(n: int = 1 << 20, seed: int = 1, sh_degree: int = 3, adjacency: bool = True,
                angular_sigma: float = 0.5) -> HostScene:
    """Mixture-on-shell scene; density LogNormal(0, 1.5); SH coefficients N(0, 0.3)."""
    rng = np.random.default_rng(seed)
    xyz = shell_points(n, rng, angular_sigma)
    density = np.exp(rng.normal(scale=1.5, size=n)).astype(np.float32)
    K = (sh_degree + 1) ** 2
    sh = rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32)
    if adjacency:
        off, nb = delaunay_csr(xyz)
    else:
        off, nb = np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32)
    return HostScene(xyz, density, sh, sh_degree, off, nb)


def power_scene(n: int, seed: int = 4, adjacency: bool = True) -> HostScene:
    """Relighting scene of power cells: the mixture-on-shell sites, radii
    log-uniform in [1e-3, 1e-1], density LogNormal(0, 1) scaled by 20 (the
    cells are small, so a visible surface needs dense cells), diffuse albedo
    U[0,1]^3, GGX roughness U[0.1, 0.9], SH degree 0 (colour comes from the
    lights).  Adjacency is the Delaunay graph of the sites."""
    rng = np.random.default_rng(seed)
    xyz = shell_points(n, rng)
    radius = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), n)).astype(np.float32)
    density = (20.0 * np.exp(rng.normal(size=n))).astype(np.float32)
    material = np.concatenate([rng.uniform(0, 1, size=(n, 3)), rng.uniform(0.1, 0.9, size=(n, 1))], 1).astype(np.float32)
    sh = np.zeros((n, 3), np.float32)
    if adjacency:
        off, nb = delaunay_csr(xyz)
    else:
        off, nb = np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32)
    return HostScene(xyz, density, sh, 0, off, nb, radius, material)


def cech_csr(xyz: np.ndarray, radius: np.ndarray, bands: int = 8, chunk: int = 1 << 18) -> tuple:
    """Cech adjacency of weighted sites (adjacency.rs:406-540): an edge {i, j}
    whenever |p_i - p_j|^2 <= (r_i + r_j)^2, evaluated in FP32 as
    ((dx*dx) + (dy*dy)) + (dz*dz) against (r_i + r_j)^2; rows sorted and
    de-duplicated (adjacency.rs:60-84).  Negative and NaN radii count as
    zero (Rust's f32::max(r, 0.0) returns 0.0 for NaN).  Sites
    are split into logarithmic radius bands (as the reference does) so each
    band pair is searched to its own maximum distance with a k-d tree."""
    from scipy.spatial import cKDTree

    xyz = np.ascontiguousarray(xyz, np.float32)
    r = np.fmax(np.ascontiguousarray(radius, np.float32), np.float32(0))
    n = len(r)
    pos = r[r > 0]
    rmin = float(pos.min()) if len(pos) else 0.0
    rmax = float(r.max()) if n else 0.0
    if n == 0:
        return np.zeros(1, np.uint32), np.zeros(0, np.uint32)
    if rmax > rmin > 0:
        band = np.floor(np.log(np.maximum(r, rmin) / rmin) / math.log(rmax / rmin) * bands).astype(np.int64)
        band = np.clip(band, 0, bands - 1)
    else:
        band = np.zeros(n, np.int64)
    members = [np.nonzero(band == b)[0] for b in range(bands)]
    trees = [cKDTree(xyz[m].astype(np.float64)) if len(m) else None for m in members]
    bmax = [float(r[m].max()) if len(m) else 0.0 for m in members]
    src, dst = [], []
    for a in range(bands):
        for b in range(a, bands):
            if trees[a] is None or trees[b] is None:
                continue
            reach = (bmax[a] + bmax[b]) * (1 + 1e-5) + 1e-7
            ma, mb = members[a], members[b]
            for c0 in range(0, len(ma), chunk):
                qa = ma[c0:c0 + chunk]
                sub = cKDTree(xyz[qa].astype(np.float64))
                m = sub.sparse_distance_matrix(trees[b], reach, output_type="ndarray")
                i = qa[m["i"]]
                j = mb[m["j"]]
                keep = i != j
                if a == b:
                    keep &= i < j
                i, j = i[keep], j[keep]
                d = xyz[i] - xyz[j]
                d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
                s_ = r[i] + r[j]
                ok = d2 <= s_ * s_
                src.append(i[ok]); dst.append(j[ok])
    i = np.concatenate(src) if src else np.zeros(0, np.int64)
    j = np.concatenate(dst) if dst else np.zeros(0, np.int64)
    rows = np.concatenate([i, j])
    cols = np.concatenate([j, i])
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    if len(rows):
        dup = np.zeros(len(rows), bool)
        dup[1:] = (rows[1:] == rows[:-1]) & (cols[1:] == cols[:-1])
        rows, cols = rows[~dup], cols[~dup]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, rows + 1, 1)
    return np.cumsum(off).astype(np.uint32), cols.astype(np.uint32)


def powerfoam_scene(n: int, seed: int = 4, sh_degree: int = 0, density_scale: float = 20.0,
                    adjacency: bool = True, cech=None) -> HostScene:
    """PowerFoam scene of the serving workload: n sites uniform in a cube of
    side 2 (n / 65536)^(1/3) centred at the origin (65,536 sites fill [-1,1]^3;
    larger scenes cover more space at the same site density), radii
    log-uniform in [1e-3, 1e-1], density LogNormal(0, 1) scaled by
    density_scale, diffuse albedo U[0,1]^3, GGX roughness U[0.1, 0.9], SH
    coefficients N(0, 0.3) of the given degree.  Adjacency is the Cech graph
    of the support spheres (what the splat tracer clips against), built by
    `cech(xyz, radius) -> (offsets, neighbors)` (default cech_csr; the
    device builder bv.compute_cech gives the same rows)."""
    rng = np.random.default_rng(seed)
    half = (n / 65536.0) ** (1.0 / 3.0)
    xyz = rng.uniform(-half, half, size=(n, 3)).astype(np.float32)
    radius = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), n)).astype(np.float32)
    density = (density_scale * np.exp(rng.normal(size=n))).astype(np.float32)
    material = np.concatenate([rng.uniform(0, 1, size=(n, 3)), rng.uniform(0.1, 0.9, size=(n, 1))], 1).astype(np.float32)
    K = (sh_degree + 1) ** 2
    sh = rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32)
    if adjacency:
        off, nb = (cech or cech_csr)(xyz, radius)
    else:
        off, nb = np.zeros(n + 1, np.uint32), np.zeros(0, np.uint32)
    return HostScene(xyz, density, sh, sh_degree, off, nb, radius, material)


@dataclass
class GaussianHostScene:
    """A Gaussian PointCloudModel: mean [N,3], opacity [N], rotation [N,4]
    quaternion (x, y, z, w), scale [N,3], SH [N,3K] (RGB interleaved)."""
    mean: np.ndarray
    opacity: np.ndarray
    rotation: np.ndarray
    scale: np.ndarray
    sh: np.ndarray
    sh_degree: int

    @property
    def n(self) -> int:
        return int(self.opacity.shape[0])


def random_quaternions(n: int, rng) -> np.ndarray:
    """Uniformly random unit quaternions (Shoemake's method), (x, y, z, w)."""
    u1, u2, u3 = rng.uniform(size=(3, n))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3),
                  b * np.cos(2 * np.pi * u3)], 1)
    return q.astype(np.float32)


def gaussian_scene(n: int = 524_288, seed: int = 3, sh_degree: int = 3) -> GaussianHostScene:
    """Gaussian workload: means from the mixture-on-shell distribution, scales
    log-uniform in [1e-3, 5e-2] per axis, uniformly random unit quaternions,
    opacity sigmoid(N(0, 1)), SH coefficients N(0, 0.3)."""
    rng = np.random.default_rng(seed)
    mean = shell_points(n, rng)
    scale = np.exp(rng.uniform(np.log(1e-3), np.log(5e-2), size=(n, 3))).astype(np.float32)
    rot = random_quaternions(n, rng)
    opacity = (1.0 / (1.0 + np.exp(-rng.normal(size=n)))).astype(np.float32)
    K = (sh_degree + 1) ** 2
    sh = rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32)
    return GaussianHostScene(mean, opacity, rot, scale, sh, sh_degree)


def random_lights(count: int, rng, radius: float = 3.0) -> list:
    """count point lights on a sphere of the given radius around the origin,
    intensity U[4, 16] per channel, half of them with an angular falloff
    (exponent 2) aimed at the origin."""
    out = []
    for k in range(count):
        v = rng.normal(size=3)
        pos = radius * v / np.linalg.norm(v)
        expo = 2.0 if k % 2 else 0.0
        out.append(dict(position=pos.astype(np.float32), direction=(-pos / np.linalg.norm(pos)).astype(np.float32),
                        intensity=rng.uniform(4, 16, 3).astype(np.float32), exponent=np.float32(expo)))
    return out


def perturbed(scene: HostScene, seed: int = 2) -> HostScene:
    """Target scene for the training workload: density * exp(N(0, 0.1)),
    SH + N(0, 0.05), positions + N(0, 1e-3); adjacency kept."""
    rng = np.random.default_rng(seed)
    n = scene.n
    return HostScene((scene.xyz + rng.normal(scale=1e-3, size=(n, 3))).astype(np.float32),
                     (scene.density * np.exp(rng.normal(scale=0.1, size=n))).astype(np.float32),
                     (scene.sh + rng.normal(scale=0.05, size=scene.sh.shape)).astype(np.float32), scene.sh_degree,
                     scene.offsets, scene.neighbors, scene.radius)


def random_rays(cameras: List[Camera], width: int, height: int, count: int, rng) -> tuple:
    """`count` rays, uniform over camera and pixel (pixel centres).  Returns
    (camera index [count] u32, pixel x [count] u32, pixel y [count] u32)."""
    cam = rng.integers(0, len(cameras), count).astype(np.uint32)
    px = rng.integers(0, width, count).astype(np.uint32)
    py = rng.integers(0, height, count).astype(np.uint32)
    return cam, px, py


def containing_cell(scene: HostScene, position) -> int:
    """lib.rs:526-541 in f32: argmin |x - p|^2 - r^2, first index on ties."""
    q = np.asarray(position, np.float32)
    d = scene.xyz - q[None, :]
    dist = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    if scene.radius is not None:
        dist = dist - scene.radius * scene.radius
    return int(np.argmin(dist))
