"""RadFoam / PowerFoam PLY reader and writer (ASCII and binary little-endian).

Reader: blade-volume/src/io/radfoam_ply.rs.
* Header (`read_header`, :452-597): magic `ply`, one `format ascii 1.0` or
  `format binary_little_endian 1.0` line, `comment` / `obj_info` lines,
  scalar properties float / float32, uint / uint32, uchar / uint8 (no lists),
  no duplicate elements or properties, `end_header`.
* Schema (`Schema::new`, :123-372): elements exactly `vertex` then
  `adjacency`; vertex x, y, z, density (float) and adjacency_offset (uint, the
  CSR end offset of the row); DC from blade_sh_dc_0..2 (float, exact) or else
  from the red/green/blue preview (uchar), DC = (rgb8 / 255 - 0.5) / C0
  (`set_preview_dc`, :722-726); color_sh_k (float) -> coefficient slot 3 + k,
  with the SH degree inferred from their count (`infer_sh_degree`, :435-450);
  optional radius (float, PowerFoam); adjacency carries one uint per entry.
* Body: binary rows of the packed property stride, whose total must equal the
  file's body size (`try_load`, :1141-1170); ASCII rows skip blank and
  `comment` lines (`read_ascii_data_line`, :861-882).
* CSR checks (`validate_csr`, :1111-1139) and finite SH (lib.rs:271-273).

Oriented-surface, surface-detail, surface-colour and spherical-Voronoi
properties are rejected: the device scene does not carry those models.

Writer: blade-volume-convert/src/lib.rs:1873-2081 (`write_radfoam_ply_ascii`,
`write_radfoam_ply_binary`): the same vertex property order, exact DC beside
a preview colour round(clamp(C0 * dc + 0.5, 0, 1) * 255) (`sh0_to_color`,
:2302-2320), color_sh_j = coefficient slot 3 + j.  ASCII floats use the
shortest decimal that reads back to the same FP32 value (Rust's `Display`
for f32), so both formats round-trip bit for bit.
"""
from __future__ import annotations

import os

import numpy as np

from .scenes import SH_C0, HostScene

_SCALARS = {"float": "<f4", "float32": "<f4", "uint": "<u4", "uint32": "<u4", "uchar": "u1", "uint8": "u1"}
_UNSUPPORTED = ("nx", "ny", "nz", "surface_offset")
_UNSUPPORTED_PREFIXES = ("blade_surface_detail_", "blade_surface_color_", "blade_spherical_voronoi_")
MAX_HEADER_BYTES = 1024 * 1024  # radfoam_ply.rs:12


class PlyError(ValueError):
    """Invalid or unsupported RadFoam PLY data (the reference's LoadError::InvalidData)."""


def _read_header(f):
    fmt = None
    elements = []  # [name, count, [(prop, dtype)]]
    saw_magic = False
    used = 0
    while True:
        raw = f.readline(MAX_HEADER_BYTES + 1 - used)
        if not raw:
            raise PlyError("RadFoam PLY header has no end_header")
        used += len(raw)
        if used > MAX_HEADER_BYTES:
            raise PlyError("RadFoam PLY header exceeds 1 MiB")
        words = raw.decode("ascii", errors="replace").split()
        if not words:
            continue
        head = words[0]
        if head == "ply":
            if saw_magic or used != len(raw):
                raise PlyError("unexpected duplicate PLY magic")
            saw_magic = True
        elif head == "format":
            if fmt is not None:
                raise PlyError("duplicate RadFoam PLY format line")
            if len(words) < 3 or words[2] != "1.0":
                raise PlyError("unsupported RadFoam PLY version")
            if words[1] not in ("ascii", "binary_little_endian"):
                raise PlyError(f"unsupported RadFoam PLY format '{words[1]}'")
            fmt = words[1]
        elif head in ("comment", "obj_info"):
            continue
        elif head == "element":
            if len(words) < 3:
                raise PlyError("RadFoam PLY element needs a name and a count")
            if any(e[0] == words[1] for e in elements):
                raise PlyError(f"duplicate RadFoam PLY element '{words[1]}'")
            try:
                count = int(words[2])
            except ValueError:
                raise PlyError(f"invalid count for element '{words[1]}'") from None
            if count < 0:
                raise PlyError(f"invalid count for element '{words[1]}'")
            elements.append([words[1], count, []])
        elif head == "property":
            if not elements:
                raise PlyError("RadFoam PLY property appears before an element")
            if len(words) > 1 and words[1] == "list":
                raise PlyError(f"RadFoam PLY list properties are unsupported in element '{elements[-1][0]}'")
            if len(words) < 3:
                raise PlyError("RadFoam PLY property needs a type and a name")
            if words[1] not in _SCALARS:
                raise PlyError(f"unsupported RadFoam PLY scalar type '{words[1]}'")
            props = elements[-1][2]
            if any(p == words[2] for p, _ in props):
                raise PlyError(f"duplicate RadFoam PLY property '{elements[-1][0]}.{words[2]}'")
            props.append((words[2], np.dtype(_SCALARS[words[1]])))
        elif head == "end_header":
            break
        else:
            raise PlyError(f"unexpected RadFoam PLY header token '{head}'")
    if not saw_magic:
        raise PlyError("RadFoam PLY is missing the 'ply' magic")
    if fmt is None:
        raise PlyError("RadFoam PLY has no format line")
    return fmt, elements


def _schema(elements):
    if len(elements) != 2 or elements[0][0] != "vertex" or elements[1][0] != "adjacency":
        raise PlyError("RadFoam PLY elements must be exactly 'vertex' followed by 'adjacency'")
    (_, n, vprops), (_, e, aprops) = elements
    if n > 0xFFFFFFFF or e > 0xFFFFFFFF:
        raise PlyError("RadFoam PLY point and adjacency counts must fit in uint32")
    types = dict(vprops)
    f4, u4, u1 = np.dtype("<f4"), np.dtype("<u4"), np.dtype("u1")

    def need(name, ty):
        if name not in types:
            raise PlyError(f"RadFoam PLY element 'vertex' is missing property '{name}'")
        if types[name] != ty:
            raise PlyError(f"RadFoam PLY property 'vertex.{name}' has the wrong type")

    def triplet(names, ty):
        present = [m in types for m in names]
        if not any(present):
            return None
        if not all(present):
            raise PlyError(f"RadFoam PLY needs all of {names} or none")
        for m in names:
            need(m, ty)
        return names

    for m in ("x", "y", "z", "density"):
        need(m, f4)
    need("adjacency_offset", u4)
    exact = triplet(("blade_sh_dc_0", "blade_sh_dc_1", "blade_sh_dc_2"), f4)
    preview = triplet(("red", "green", "blue"), u1)
    if exact is None and preview is None:
        raise PlyError("RadFoam PLY vertex needs exact blade_sh_dc_0..2 or red/green/blue preview DC")
    rest = {}
    for name, ty in vprops:
        if name in _UNSUPPORTED or name.startswith(_UNSUPPORTED_PREFIXES):
            raise PlyError(f"RadFoam PLY property '{name}' (oriented / detail / surface-colour / "
                           "spherical-Voronoi models) is not supported by the device scene")
        if name.startswith("color_sh_"):
            if ty != f4:
                raise PlyError(f"RadFoam PLY property '{name}' must be float32")
            try:
                rest[int(name[len("color_sh_"):])] = name
            except ValueError:
                raise PlyError(f"invalid RadFoam PLY property '{name}'") from None
    if sorted(rest) != list(range(len(rest))):
        raise PlyError("RadFoam PLY color_sh_* properties are not contiguous from 0")
    if len(rest) % 3:
        raise PlyError(f"RadFoam PLY color_sh property count {len(rest)} is not divisible by three")
    K = len(rest) // 3 + 1
    deg = next((d for d in range(4) if (d + 1) ** 2 == K), None)
    if deg is None:
        raise PlyError(f"RadFoam PLY has {K} SH components, expected a complete degree up to 3")
    radius = None
    if "radius" in types:
        need("radius", f4)
        radius = "radius"
    if len(aprops) < 1 or "adjacency" not in dict(aprops):
        raise PlyError("RadFoam PLY element 'adjacency' is missing property 'adjacency'")
    if dict(aprops)["adjacency"] != u4:
        raise PlyError("RadFoam PLY property 'adjacency.adjacency' has the wrong type")
    return dict(n=n, e=e, vprops=vprops, aprops=aprops, exact=exact, preview=preview,
                rest=[rest[k] for k in range(len(rest))], deg=deg, K=K, radius=radius)


def _row_dtype(props):
    names = [p for p, _ in props]
    fmts = [t for _, t in props]
    offs = np.cumsum([0] + [t.itemsize for t in fmts[:-1]]).tolist()
    return np.dtype({"names": names, "formats": fmts, "offsets": offs,
                     "itemsize": int(sum(t.itemsize for t in fmts))})


def _parse_f32(tokens) -> np.ndarray:
    """Decimal tokens to correctly rounded FP32 values (Rust's
    `str::parse::<f32>`).  A token goes through the nearest double first; that
    double rounds to FP32 correctly unless it sits exactly halfway between two
    FP32 values while the token itself does not (the first rounding landed on
    the midpoint).  Those rare tokens are decided by exact rational comparison
    with the midpoint."""
    d = np.asarray(tokens, dtype=np.float64)
    with np.errstate(over="ignore"):
        f = d.astype(np.float32)
        lo = np.nextafter(f, np.float32(-np.inf)).astype(np.float64)
        hi = np.nextafter(f, np.float32(np.inf)).astype(np.float64)
    fin = np.isfinite(f)
    f64 = f.astype(np.float64)
    mid_lo, mid_hi = (f64 + lo) / 2, (f64 + hi) / 2
    hazard = np.nonzero(fin & ((d == mid_lo) | (d == mid_hi)))[0]
    if hazard.size:
        from fractions import Fraction

        for i in hazard.tolist():
            exact = Fraction(str(tokens[i]))
            mid = Fraction(float(d[i]))
            if exact == mid:
                continue  # a true tie: round-half-even, as the cast did
            below, above = (lo[i], f64[i]) if d[i] == mid_lo[i] else (f64[i], hi[i])
            f[i] = np.float32(above if exact > mid else below)
    return f


def _ascii_data_line(f):
    """Next line that is neither blank nor a comment, stripped; None at EOF."""
    while True:
        raw = f.readline()
        if not raw:
            return None
        line = raw.decode("ascii", errors="replace").strip()
        if line and not line.startswith("comment"):
            return line


def _ascii_rows(f, count, props, what):
    """`count` data rows of len(props) tokens each, as a structured array.  Rows
    are gathered first and converted one column at a time; floats are
    correctly rounded to FP32 (_parse_f32) and integers are range-checked."""
    rows = []
    while len(rows) < count:
        raw = f.readline()
        if not raw:
            raise PlyError(f"RadFoam PLY {what} body is truncated")
        line = raw.decode("ascii", errors="replace").strip()
        if not line or line.startswith("comment"):
            continue
        tok = line.split()
        if len(tok) < len(props):
            raise PlyError(f"RadFoam PLY {what} row {len(rows)} has {len(tok)} values, expected {len(props)}")
        rows.append(tok[:len(props)])
    out = np.zeros(count, _row_dtype(props))
    if count == 0:
        return out
    table = np.asarray(rows, dtype=object)
    for k, (name, ty) in enumerate(props):
        col = table[:, k]
        try:
            if ty.kind == "f":
                out[name] = _parse_f32(col)
            else:
                v = np.asarray([int(t) for t in col], dtype=np.int64)
                if v.min() < 0 or v.max() > np.iinfo(ty).max:
                    raise ValueError
                out[name] = v
        except (ValueError, OverflowError):
            raise PlyError(f"invalid RadFoam PLY value for '{name}'") from None
    return out


def load_radfoam_ply(path: str) -> HostScene:
    """PointCloudModel of a RadFoam / PowerFoam PLY file (`radfoam_ply::try_load`,
    radfoam_ply.rs:1141-1215); raises PlyError where the reference returns
    LoadError::InvalidData."""
    with open(path, "rb") as f:
        fmt, elements = _read_header(f)
        s = _schema(elements)
        n, e = s["n"], s["e"]
        vdt, adt = _row_dtype(s["vprops"]), _row_dtype(s["aprops"])
        body = os.fstat(f.fileno()).st_size - f.tell()
        if fmt == "binary_little_endian":
            expected = n * vdt.itemsize + e * adt.itemsize
            if body != expected:
                raise PlyError(f"RadFoam PLY binary body is {body} bytes, expected {expected}")
            vrows = np.frombuffer(f.read(n * vdt.itemsize), vdt, count=n)
            arows = np.frombuffer(f.read(e * adt.itemsize), adt, count=e)
        else:
            if n + e > body:
                raise PlyError("RadFoam PLY ASCII element counts are implausible for the file size")
            vrows = _ascii_rows(f, n, s["vprops"], "vertex")
            arows = _ascii_rows(f, e, s["aprops"], "adjacency")
            # radfoam_ply.rs:1101-1105
            if _ascii_data_line(f) is not None:
                raise PlyError("unexpected trailing RadFoam PLY ASCII data")
    K = s["K"]
    xyz = np.stack([vrows["x"], vrows["y"], vrows["z"]], axis=1).astype(np.float32)
    density = np.ascontiguousarray(vrows["density"], np.float32)
    sh = np.zeros((n, 3 * K), np.float32)
    if s["exact"] is not None:
        for ch, name in enumerate(s["exact"]):
            sh[:, ch] = vrows[name]
    else:
        for ch, name in enumerate(s["preview"]):
            sh[:, ch] = (vrows[name].astype(np.float32) / np.float32(255.0) - np.float32(0.5)) / np.float32(SH_C0)
    for k, name in enumerate(s["rest"]):
        sh[:, 3 + k] = vrows[name]
    offsets = np.zeros(n + 1, np.uint32)
    offsets[1:] = vrows["adjacency_offset"]
    neighbors = np.ascontiguousarray(arows["adjacency"], np.uint32)
    radius = np.ascontiguousarray(vrows[s["radius"]], np.float32) if s["radius"] else None
    # validate_csr (radfoam_ply.rs:1111-1139)
    if int(offsets[n]) != e:
        raise PlyError(f"RadFoam PLY final CSR offset is {int(offsets[n])}, expected {e}")
    if n and np.any(np.diff(offsets.astype(np.int64)) < 0):
        raise PlyError("invalid RadFoam PLY CSR range")
    if e and int(neighbors.max()) >= n:
        bad = int(np.argmax(neighbors >= n))
        raise PlyError(f"RadFoam PLY adjacency entry {bad} references point {int(neighbors[bad])}, "
                       f"but there are {n} points")
    if not np.all(np.isfinite(sh)):
        raise PlyError("RadFoam PLY model: SH coefficients must be finite")
    return HostScene(xyz, density, sh, s["deg"], offsets, neighbors, radius)


# The tests and older callers name the ASCII case explicitly.
load_radfoam_ascii = load_radfoam_ply


def _f32_text(v) -> str:
    """Shortest decimal that parses back to the same FP32 value, without an
    exponent (Rust's `Display` for f32)."""
    return np.format_float_positional(np.float32(v), unique=True, trim="-")


def preview_rgb8(sh_dc: np.ndarray) -> np.ndarray:
    """[n, 3] uchar preview of the DC coefficients: round(clamp(C0 * dc + 0.5,
    0, 1) * 255), rounding halves away from zero (`sh0_to_color`,
    blade-volume-convert/src/lib.rs:2302-2320, and its caller's `round()`)."""
    c = np.clip(sh_dc.astype(np.float32) * np.float32(SH_C0) + np.float32(0.5), np.float32(0), np.float32(1))
    x = c * np.float32(255.0)
    lo = np.floor(x)
    r = np.where(x - lo >= np.float32(0.5), lo + 1, lo)
    return np.clip(r, 0, 255).astype(np.uint8)


def save_radfoam_ply(path: str, scene: HostScene, binary: bool = True) -> None:
    """Write `scene` as a RadFoam PLY (PowerFoam when it has radii) with the
    property order of write_radfoam_ply_ascii / write_radfoam_ply_binary
    (blade-volume-convert/src/lib.rs:1873-2081)."""
    n = scene.n
    K = scene.sh.shape[1] // 3
    n_rest = 3 * (K - 1)
    e = int(scene.neighbors.shape[0])
    if scene.offsets.shape != (n + 1,) or int(scene.offsets[0]) != 0 or int(scene.offsets[n]) != e:
        raise PlyError("scene CSR offsets do not describe its neighbour array")
    props = [("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("density", "<f4"), ("adjacency_offset", "<u4"),
             ("red", "u1"), ("green", "u1"), ("blue", "u1"),
             ("blade_sh_dc_0", "<f4"), ("blade_sh_dc_1", "<f4"), ("blade_sh_dc_2", "<f4")]
    props += [(f"color_sh_{j}", "<f4") for j in range(n_rest)]
    if scene.radius is not None:
        props.append(("radius", "<f4"))
    head = ["ply", "format binary_little_endian 1.0" if binary else "format ascii 1.0", f"element vertex {n}"]
    head += [f"property {'float' if t == '<f4' else ('uint' if t == '<u4' else 'uchar')} {p}" for p, t in props]
    head += [f"element adjacency {e}", "property uint adjacency", "end_header"]
    rows = np.zeros(n, _row_dtype([(p, np.dtype(t)) for p, t in props]))
    rows["x"], rows["y"], rows["z"] = scene.xyz[:, 0], scene.xyz[:, 1], scene.xyz[:, 2]
    rows["density"] = scene.density
    rows["adjacency_offset"] = scene.offsets[1:]
    rgb = preview_rgb8(scene.sh[:, :3])
    rows["red"], rows["green"], rows["blue"] = rgb[:, 0], rgb[:, 1], rgb[:, 2]
    for ch in range(3):
        rows[f"blade_sh_dc_{ch}"] = scene.sh[:, ch]
    for j in range(n_rest):
        rows[f"color_sh_{j}"] = scene.sh[:, 3 + j]
    if scene.radius is not None:
        rows["radius"] = scene.radius
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        if binary:
            f.write(rows.tobytes())
            f.write(np.ascontiguousarray(scene.neighbors, "<u4").tobytes())
            return
        for r in range(n):
            vals = rows[r]
            f.write((" ".join(_f32_text(vals[p]) if t == "<f4" else str(int(vals[p])) for p, t in props)
                     + "\n").encode("ascii"))
        f.write("".join(f"{int(v)}\n" for v in scene.neighbors).encode("ascii"))
