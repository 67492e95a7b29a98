"""Wavefront OBJ + MTL scene input for the path tracer.

Builds the Triangle and TriangleAttrib arrays the way the reference demo
does from tinyobjloader's output (main_gpu.cpp:209-282): one Triangle per
(triangulated) face, the three face-vertex normals as shading normals, the
face material's diffuse colour (`Kd`) as reflectance and its emission (`Ke`)
as emissive; a face without a known material gets reflectance (0.5, 0.5, 0.5)
and no emission.

tinyobjloader (the reference's OBJ reader, pulled in through its un-vendored
prlib submodule, version not pinned in the reference tree) is restated here
for what that loop consumes:
  * `v x y z` positions, `vn x y z` normals; `vt` is read past;
  * `f` tokens `v`, `v/vt`, `v//vn`, `v/vt/vn`, 1-based, negative = relative
    to the end of the list so far;
  * `usemtl name` selects the material whose `newmtl` came at that position
    in the `mtllib` file(s) (unknown names select none);
  * triangulation: a quad is split along its shorter diagonal,
    (0,1,2)+(0,2,3) when |v2-v0|^2 < |v3-v1|^2, else (0,1,3)+(1,2,3), as
    tinyobjloader 2.x does; a face of more than four vertices is fanned from
    its first vertex (tinyobjloader clips ears there; the scenes used here
    have none).
A face-vertex without a normal gets a zero shading normal (the reference
reads out of bounds there).
"""
from __future__ import annotations

import gzip
import os
from dataclasses import dataclass, field

import numpy as np

ATTRIB_FLOATS = 15  # TriangleAttrib: shadingNormals[3], reflectance, emissive


@dataclass
class Material:
    name: str
    diffuse: tuple = (0.0, 0.0, 0.0)
    emission: tuple = (0.0, 0.0, 0.0)


@dataclass
class ObjScene:
    triangles: np.ndarray  # (n, 3, 3) float32
    attribs: np.ndarray  # (n, 15) float32
    materials: list = field(default_factory=list)
    material_ids: np.ndarray = None  # (n,) int32, -1 = none
    faces: int = 0  # polygons before triangulation


def parse_mtl(path: str) -> list:
    mats: list = []
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "newmtl":
                mats.append(Material(" ".join(tok[1:])))
            elif mats and tok[0] in ("Kd", "Ke") and len(tok) >= 4:
                rgb = (float(tok[1]), float(tok[2]), float(tok[3]))
                if tok[0] == "Kd":
                    mats[-1].diffuse = rgb
                else:
                    mats[-1].emission = rgb
    return mats


def _index(tok: str, count: int) -> int:
    i = int(tok)
    return i - 1 if i > 0 else count + i


def _triangulate(pos: list) -> list:
    """Corner triples (indices into the face's vertex list)."""
    k = len(pos)
    if k == 3:
        return [(0, 1, 2)]
    if k == 4:
        p = [np.asarray(x, np.float32) for x in pos]
        e02, e13 = p[2] - p[0], p[3] - p[1]
        sqr02 = np.float32(e02[0] * e02[0] + e02[1] * e02[1] + e02[2] * e02[2])
        sqr13 = np.float32(e13[0] * e13[0] + e13[1] * e13[1] + e13[2] * e13[2])
        return [(0, 1, 2), (0, 2, 3)] if sqr02 < sqr13 else [(0, 1, 3), (1, 2, 3)]
    return [(0, j, j + 1) for j in range(1, k - 1)]


def _open_text(path: str):
    if path.endswith(".gz"):
        return gzip.open(path, "rt", encoding="utf-8", errors="replace")
    return open(path, "r", encoding="utf-8", errors="replace")


def load_obj(path: str) -> ObjScene:
    """Reads an OBJ file (plain, or gzip-compressed when the name ends in
    .gz); `mtllib` files are looked up beside it."""
    verts: list = []
    normals: list = []
    materials: list = []
    mat_index: dict = {}
    current = -1
    tris, attrs, mids = [], [], []
    nfaces = 0
    base = os.path.dirname(os.path.abspath(path))
    with _open_text(path) as fh:
        for line in fh:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            key = tok[0]
            if key == "v":
                verts.append((float(tok[1]), float(tok[2]), float(tok[3])))
            elif key == "vn":
                normals.append((float(tok[1]), float(tok[2]), float(tok[3])))
            elif key == "mtllib":
                for name in tok[1:]:
                    mp = os.path.join(base, name)
                    if os.path.exists(mp):
                        for m in parse_mtl(mp):
                            if m.name not in mat_index:
                                mat_index[m.name] = len(materials)
                                materials.append(m)
            elif key == "usemtl":
                current = mat_index.get(" ".join(tok[1:]), -1)
            elif key == "f":
                corners = []
                for t in tok[1:]:
                    parts = t.split("/")
                    vi = _index(parts[0], len(verts))
                    ni = _index(parts[2], len(normals)) if len(parts) > 2 and parts[2] else -1
                    corners.append((vi, ni))
                if len(corners) < 3:
                    continue
                nfaces += 1
                pos = [verts[c[0]] for c in corners]
                for a, b, c in _triangulate(pos):
                    tri = [pos[a], pos[b], pos[c]]
                    sn = [normals[corners[x][1]] if corners[x][1] >= 0 else (0.0, 0.0, 0.0) for x in (a, b, c)]
                    if current >= 0:
                        refl, emis = materials[current].diffuse, materials[current].emission
                    else:
                        refl, emis = (0.5, 0.5, 0.5), (0.0, 0.0, 0.0)
                    tris.append(tri)
                    attrs.append([*sn[0], *sn[1], *sn[2], *refl, *emis])
                    mids.append(current)
    triangles = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    attribs = np.asarray(attrs, np.float32).reshape(-1, ATTRIB_FLOATS)
    return ObjScene(triangles, attribs, materials, np.asarray(mids, np.int32), nfaces)


def cornell_box_obj(directory: str) -> str:
    """Write a Cornell box as OBJ + MTL (the classic room of five walls, a
    short and a tall box and an emissive ceiling quad, in quads with per-face
    normals) into `directory`; returns the OBJ path."""
    mtl = {
        "white": ((0.725, 0.71, 0.68), (0, 0, 0)),
        "red": ((0.63, 0.065, 0.05), (0, 0, 0)),
        "green": ((0.14, 0.45, 0.091), (0, 0, 0)),
        "light": ((0.78, 0.78, 0.78), (17.0, 12.0, 4.0)),
    }
    quads = [  # (material, 4 corners counter-clockwise seen from inside)
        ("white", [(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)]),  # floor
        ("white", [(-1, 2, -1), (-1, 2, 1), (1, 2, 1), (1, 2, -1)]),  # ceiling
        ("white", [(-1, 0, -1), (-1, 2, -1), (1, 2, -1), (1, 0, -1)]),  # back
        ("red", [(-1, 0, -1), (-1, 0, 1), (-1, 2, 1), (-1, 2, -1)]),  # left
        ("green", [(1, 0, -1), (1, 2, -1), (1, 2, 1), (1, 0, 1)]),  # right
        ("light", [(-0.25, 1.99, -0.25), (0.25, 1.99, -0.25), (0.25, 1.99, 0.25), (-0.25, 1.99, 0.25)]),
    ]

    def box(cx, cz, sx, sz, hgt, rot):
        c, s = np.cos(rot), np.sin(rot)
        pts = []
        for dx, dz in ((-sx, -sz), (sx, -sz), (sx, sz), (-sx, sz)):
            pts.append((cx + c * dx - s * dz, cz + s * dx + c * dz))
        out = [("white", [(x, hgt, z) for x, z in pts])]  # top
        for i in range(4):
            (x0, z0), (x1, z1) = pts[i], pts[(i + 1) % 4]
            out.append(("white", [(x0, 0, z0), (x1, 0, z1), (x1, hgt, z1), (x0, hgt, z0)]))
        return out

    quads += box(0.33, 0.35, 0.29, 0.29, 0.6, -0.29) + box(-0.35, -0.3, 0.29, 0.29, 1.2, 0.29)
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "cornell.mtl"), "w") as fh:
        for name, (kd, ke) in mtl.items():
            fh.write(f"newmtl {name}\nKd {kd[0]} {kd[1]} {kd[2]}\nKe {ke[0]} {ke[1]} {ke[2]}\n\n")
    path = os.path.join(directory, "cornell.obj")
    with open(path, "w") as fh:
        fh.write("mtllib cornell.mtl\no CornellBox\n")
        nv = nn = 0
        for name, q in quads:
            p = np.asarray(q, np.float64)
            n = np.cross(p[1] - p[0], p[2] - p[0])
            n /= np.linalg.norm(n)
            for v in p:
                fh.write(f"v {v[0]:.6f} {v[1]:.6f} {v[2]:.6f}\n")
            fh.write(f"vn {n[0]:.6f} {n[1]:.6f} {n[2]:.6f}\n")
            fh.write(f"usemtl {name}\n")
            fh.write("f " + " ".join(f"{nv + k + 1}//{nn + 1}" for k in range(4)) + "\n")
            nv += 4
            nn += 1
    return path


# Camera of the reference demo for the Cornell box (main_gpu.cpp:160-161):
# origin (0, 1, 4) looking at (0, 1, 0), y up; its field of view comes from
# prlib's Camera3D default, which is not in the reference tree: 45 degrees.
CAMERA_CORNELL = ((0.0, 1.0, 4.0), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 45.0)
