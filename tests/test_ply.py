"""RadFoam / PowerFoam PLY reader and writer (CPU only).

Reader: blade-volume/src/io/radfoam_ply.rs; writer:
blade-volume-convert/src/lib.rs:1873-2081.  The reference's own fixture
(tests/data/radfoam_tiny_ascii.ply) and its loader test
(`malformed_and_implausible_files_return_errors_before_allocation`,
radfoam_ply.rs:1227-1244) pin the reader; round trips through both writer
formats pin the writer bit for bit, and the preview colour is checked against
values fixed by hand (C0 * 0 + 0.5 -> 127.5 -> 128, clamps at 0 and 255).
"""
import os

import numpy as np
import pytest

import oracle
from bv import io, scenes

from conftest import GOLDEN


def _scene(n=64, deg=3, radius=True, seed=5):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    K = (deg + 1) ** 2
    density = np.exp(rng.normal(size=n)).astype(np.float32)
    sh = rng.normal(scale=0.3, size=(n, 3 * K)).astype(np.float32)
    r = rng.uniform(0.05, 0.3, size=n).astype(np.float32) if radius else None
    return scenes.HostScene(xyz, density, sh, deg, off, nb, r)


def _same(a, b):
    assert a.n == b.n and a.sh_degree == b.sh_degree
    for name in ("xyz", "density", "sh", "offsets", "neighbors"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.dtype == y.dtype and x.shape == y.shape
        assert x.tobytes() == y.tobytes(), name
    assert (a.radius is None) == (b.radius is None)
    if a.radius is not None:
        assert a.radius.tobytes() == b.radius.tobytes()


@pytest.mark.parametrize("binary", [True, False])
@pytest.mark.parametrize("deg,radius", [(0, False), (1, True), (3, True), (3, False)])
def test_round_trip_is_bit_exact(tmp_path, binary, deg, radius):
    s = _scene(deg=deg, radius=radius)
    p = str(tmp_path / "scene.ply")
    io.save_radfoam_ply(p, s, binary=binary)
    _same(s, io.load_radfoam_ply(p))


def test_binary_layout_matches_the_writer_schema(tmp_path):
    # vertex row: x y z density (16) + adjacency_offset (4) + rgb (3) + exact DC (12)
    # + 45 color_sh (180) + radius (4) = 219 bytes, then one uint per adjacency entry.
    s = _scene(n=32, deg=3, radius=True)
    p = str(tmp_path / "b.ply")
    io.save_radfoam_ply(p, s, binary=True)
    raw = open(p, "rb").read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    header = raw[:end].decode().splitlines()
    assert header[:3] == ["ply", "format binary_little_endian 1.0", "element vertex 32"]
    assert header[3:8] == ["property float x", "property float y", "property float z", "property float density",
                           "property uint adjacency_offset"]
    assert header[-4:] == ["property float radius", f"element adjacency {len(s.neighbors)}",
                           "property uint adjacency", "end_header"]
    assert len(raw) - end == 32 * 219 + 4 * len(s.neighbors)
    row0 = raw[end:end + 219]
    assert np.frombuffer(row0[16:20], "<u4")[0] == s.offsets[1]
    assert np.frombuffer(row0[23:35], "<f4").tobytes() == s.sh[0, :3].tobytes()
    assert np.frombuffer(raw[-4 * len(s.neighbors):], "<u4").tobytes() == s.neighbors.tobytes()


def test_preview_colour_rounding():
    dc = np.array([[0.0, -10.0, 10.0]], np.float32)
    assert io.preview_rgb8(dc).tolist() == [[128, 0, 255]]
    # exact DC from the preview alone: (rgb8 / 255 - 0.5) / C0
    full = (np.array([1.0, 2.0, 3.0], np.float32) / np.float32(255.0) - np.float32(0.5)) / np.float32(scenes.SH_C0)
    assert io.preview_rgb8(full[None]).tolist() == [[1, 2, 3]]


def test_preview_only_binary_file(tmp_path):
    # The reference's loader takes the DC from red/green/blue when no exact DC is stored.
    p = str(tmp_path / "preview.ply")
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 2\nproperty float x\nproperty float y\n"
            "property float z\nproperty float density\nproperty uint adjacency_offset\nproperty uchar red\n"
            "property uchar green\nproperty uchar blue\nelement adjacency 2\nproperty uint adjacency\nend_header\n")
    dt = np.dtype({"names": ["x", "y", "z", "d", "o", "r", "g", "b"],
                   "formats": ["<f4"] * 4 + ["<u4", "u1", "u1", "u1"],
                   "offsets": [0, 4, 8, 12, 16, 20, 21, 22], "itemsize": 23})
    rows = np.zeros(2, dt)
    rows["x"] = [0, 1]
    rows["d"] = [0.5, 2.0]
    rows["o"] = [1, 2]
    rows["r"], rows["g"], rows["b"] = [0, 255], [128, 64], [7, 200]
    with open(p, "wb") as f:
        f.write(head.encode())
        f.write(rows.tobytes())
        f.write(np.array([1, 0], "<u4").tobytes())
    m = io.load_radfoam_ply(p)
    assert m.sh_degree == 0 and m.radius is None and list(m.offsets) == [0, 1, 2]
    want = (np.array([[0, 128, 7], [255, 64, 200]], np.float32) / np.float32(255.0) - np.float32(0.5)) \
        / np.float32(scenes.SH_C0)
    assert m.sh.tobytes() == want.astype(np.float32).tobytes()


def test_malformed_and_implausible_files_are_rejected(tmp_path):
    # radfoam_ply.rs:1227-1244
    p = str(tmp_path / "m.ply")
    open(p, "wb").write(b"ply\nformat ascii 1.0\nend_header\n")
    with pytest.raises(io.PlyError):
        io.load_radfoam_ply(p)
    open(p, "wb").write(b"ply\nformat binary_little_endian 1.0\nelement vertex 4294967295\nproperty float x\n"
                        b"property float y\nproperty float z\nproperty float density\nproperty uint adjacency_offset\n"
                        b"property uchar red\nproperty uchar green\nproperty uchar blue\nelement adjacency 0\n"
                        b"property uint adjacency\nend_header\n")
    with pytest.raises(io.PlyError, match="binary body"):
        io.load_radfoam_ply(p)


def test_invalid_bodies_and_headers_are_rejected(tmp_path):
    s = _scene(n=16, deg=1, radius=False)
    good = str(tmp_path / "g.ply")
    io.save_radfoam_ply(good, s, binary=True)
    raw = open(good, "rb").read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    p = str(tmp_path / "bad.ply")
    # truncated body
    open(p, "wb").write(raw[:-4])
    with pytest.raises(io.PlyError, match="binary body"):
        io.load_radfoam_ply(p)
    # neighbour index out of range
    bad = bytearray(raw)
    bad[-4:] = np.array([16], "<u4").tobytes()
    open(p, "wb").write(bytes(bad))
    with pytest.raises(io.PlyError, match="references point 16"):
        io.load_radfoam_ply(p)
    # list property, unsupported model properties, wrong element order
    for old, new, msg in [(b"property uint adjacency\n", b"property list uchar uint adjacency\n", "list"),
                          (b"property float density\n", b"property float density\nproperty float nx\n", "nx"),
                          (b"element adjacency", b"element edges", "elements")]:
        open(p, "wb").write(raw[:end].replace(old, new, 1) + raw[end:])
        with pytest.raises(io.PlyError, match=msg):
            io.load_radfoam_ply(p)


def test_tiny_fixture_loads_and_round_trips(tmp_path):
    m = io.load_radfoam_ply(os.path.join(GOLDEN, "radfoam_tiny_ascii.ply"))
    assert m.n == 4 and m.sh_degree == 1 and list(m.neighbors) == [1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2]
    for binary in (True, False):
        p = str(tmp_path / f"t{int(binary)}.ply")
        io.save_radfoam_ply(p, m, binary=binary)
        back = io.load_radfoam_ply(p)
        _same(m, back)
        # the written preview colour is the fixture's own (128, 128, 128)
        assert io.preview_rgb8(back.sh[:, :3]).tolist() == [[128, 128, 128]] * 4


def test_loaded_scene_traces_like_the_original(tmp_path):
    s = _scene(n=256, deg=3, radius=False, seed=9)
    p = str(tmp_path / "t.ply")
    io.save_radfoam_ply(p, s, binary=True)
    m = io.load_radfoam_ply(p)
    rng = np.random.default_rng(3)
    o = np.zeros((8, 3), np.float32)
    d = rng.normal(size=(8, 3)).astype(np.float32)
    a = oracle.trace_rays(oracle.of_host(s), o, d, int(np.argmin((s.xyz ** 2).sum(1))), weight_threshold=1e-3,
                          max_steps=256, depth=100.0)
    b = oracle.trace_rays(oracle.of_host(m), o, d, int(np.argmin((m.xyz ** 2).sum(1))), weight_threshold=1e-3,
                          max_steps=256, depth=100.0)
    assert a["rgba"].tobytes() == b["rgba"].tobytes()


def test_ascii_floats_round_once_to_fp32():
    # 1 + 2^-24 is the midpoint between FP32 1.0 and its successor.  A decimal
    # just above it rounds to the midpoint as a double, and the double's
    # round-half-even to FP32 would give 1.0; the correctly rounded FP32 value
    # (Rust's str::parse::<f32>) is the successor.
    up = np.nextafter(np.float32(1.0), np.float32(2.0))
    toks = ["1.000000059604644775390625000001", "1.000000059604644775390625", "0.999999", "-1.000000059604644775390625000001",
            "3.4028235e38", "1e-45", "0.1"]
    got = io._parse_f32(toks)
    assert got[0] == up
    assert got[1] == np.float32(1.0)  # an exact tie rounds half to even
    assert got[2] == np.float32(0.999999)
    assert got[3] == -up
    assert got[4] == np.finfo(np.float32).max
    assert got[5] == np.float32(1e-45)
    assert got[6] == np.float32(0.1)
    # every value the writer emits reads back unchanged
    rng = np.random.default_rng(2)
    vals = rng.normal(size=2000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 2000).astype(np.float32)
    vals = vals[np.isfinite(vals)]
    back = io._parse_f32([io._f32_text(v) for v in vals])
    assert back.tobytes() == vals.tobytes()


def test_trailing_ascii_data_is_rejected(tmp_path):
    # radfoam_ply.rs:1101-1105
    raw = open(os.path.join(GOLDEN, "radfoam_tiny_ascii.ply"), "rb").read()
    p = str(tmp_path / "t.ply")
    open(p, "wb").write(raw.rstrip(b"\n") + b"\n\ncomment trailing comment\n\n")
    assert io.load_radfoam_ply(p).n == 4
    open(p, "wb").write(raw.rstrip(b"\n") + b"\n7\n")
    with pytest.raises(io.PlyError, match="unexpected trailing RadFoam PLY ASCII data"):
        io.load_radfoam_ply(p)
