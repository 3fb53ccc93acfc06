"""RadFoam PLY reader (ASCII subset) for the reference's scene files.

Follows blade-volume/src/io/radfoam_ply.rs: vertex properties x, y, z,
density, adjacency_offset (CSR end offset of the row), preview colour
red/green/blue (DC = (rgb8 / 255 - 0.5) / C0 when no exact DC is stored,
:752-760), color_sh_k -> coefficient slot 3 + k (:762-764), optional radius;
element adjacency carries the flattened neighbour list.  SH degree follows from
the number of color_sh_* properties (:435-445).
"""
from __future__ import annotations

import numpy as np

from .scenes import SH_C0, HostScene


def load_radfoam_ascii(path: str) -> HostScene:
    with open(path, "r") as f:
        lines = [ln.strip() for ln in f]
    if not lines or lines[0] != "ply":
        raise ValueError("not a PLY file")
    if "format ascii 1.0" not in lines[1]:
        raise ValueError("only ASCII RadFoam PLY is supported by this reader")
    elements = []  # (name, count, [props])
    i = 2
    while lines[i] != "end_header":
        tok = lines[i].split()
        if tok and tok[0] == "element":
            elements.append((tok[1], int(tok[2]), []))
        elif tok and tok[0] == "property":
            elements[-1][2].append(tok[-1])
        i += 1
    i += 1
    rows = [ln for ln in lines[i:] if ln and not ln.startswith("comment")]
    data = {}
    pos = 0
    for name, count, props in elements:
        vals = [ln.split() for ln in rows[pos:pos + count]]
        pos += count
        data[name] = (props, vals)
    vprops, vrows = data["vertex"]
    col = {p: k for k, p in enumerate(vprops)}
    n = len(vrows)
    rest = sorted(int(p[len("color_sh_"):]) for p in vprops if p.startswith("color_sh_"))
    if rest != list(range(len(rest))) or len(rest) % 3:
        raise ValueError("color_sh_* properties must be contiguous and a multiple of three")
    K = len(rest) // 3 + 1
    deg = int(round(np.sqrt(K))) - 1
    if (deg + 1) ** 2 != K:
        raise ValueError("color_sh_* count does not match an SH degree")
    xyz = np.zeros((n, 3), np.float32)
    density = np.zeros(n, np.float32)
    sh = np.zeros((n, 3 * K), np.float32)
    offsets = np.zeros(n + 1, np.uint32)
    radius = np.zeros(n, np.float32) if "radius" in col else None
    for r, v in enumerate(vrows):
        xyz[r] = [np.float32(v[col["x"]]), np.float32(v[col["y"]]), np.float32(v[col["z"]])]
        density[r] = np.float32(v[col["density"]])
        offsets[r + 1] = np.uint32(int(v[col["adjacency_offset"]]))
        if all(c in col for c in ("red", "green", "blue")):
            for ch, c in enumerate(("red", "green", "blue")):
                sh[r, ch] = (np.float32(int(v[col[c]])) / np.float32(255.0) - np.float32(0.5)) / np.float32(SH_C0)
        for k in rest:
            sh[r, 3 + k] = np.float32(v[col[f"color_sh_{k}"]])
        if radius is not None:
            radius[r] = np.float32(v[col["radius"]])
    _, arows = data.get("adjacency", ([], []))
    neighbors = np.asarray([int(a[0]) for a in arows], np.uint32)
    return HostScene(xyz, density, sh, deg, offsets, neighbors, radius)
